// BAM alignment records on the device (gather by the sort's permutation, the
// duplicate flag applied in the same pass) and BGZF framing on the host.
//
// Record format: _pack_bam_aln, fq2bam-meth/src/linear_gpu_fm.mojo:497-665;
// gather + 0x400: _gather_perm_to_writer (:177-225) and _or_dup_flag (:408-420);
// SE field derivation: _emit_bam_from_batch (:755-800); BGZF members:
// _bgzf_block / _BgzfWriter, fq2bam-meth/python/bam_emit.py:77-185.
//
// Design: bam_layout_kernel reads the records in source order (the field
// arrays in sequence) and writes each one's 64-byte layout (header and CIGAR
// words, name / row / flags) to its output position inv[i] (inv = the
// permutation's inverse); the lengths in output order are scanned into byte
// offsets (the total lands in rec_off[n]); in bam_pack_kernel each wave takes
// 64 consecutive layouts into LDS and writes the records in turn: a record is
// assembled in LDS at the output's alignment, section by section (a byte per
// lane: header, name, CIGAR, bases from 4-bit codes, qualities, tags), then
// leaves as aligned 4-byte words (byte stores only where a word is shared with
// a neighbouring record), the next record's name, bases and qualities loading
// meanwhile, so the stream is written once, in order, with no host gather.
#include <zlib.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "ma_pe.hpp"
#include "ma_sort.hpp"

using namespace ma;

namespace {

constexpr int BAM_RG_MAX = 240;   // read-group id bytes carried in the kernel arguments
constexpr int BAM_WPB = 4;        // waves per 256-thread block
constexpr int STAGE = 256;        // name / bases / qualities bytes staged in LDS per record
constexpr int64_t BGZF_BLOCK = 65280;  // uncompressed bytes per BGZF member (bam_emit.py:25)

struct RgArg {
    int32_t n;
    uint8_t b[BAM_RG_MAX];
};

// Where a record's fields come from: SE mode derives them from merged map
// records (+ contig -> target table); otherwise ten int32 arrays (the last
// three may be null: mate tid -1, mate pos -1, tlen 0).
// se: 0 = explicit field arrays, 1 = SE merged records, 2 = PE merged records
// (2 np of them, R1 block then R2 block; output-order record i is mate i & 1 of
// pair i >> 1 and reads row (i & 1) ? np + (i >> 1) : i >> 1).
struct FieldSrc {
    int32_t se;
    int32_t n_tid;
    int64_t np;
    const uint32_t *lens;
    const int32_t *cid;
    const uint32_t *pos, *mapq;
    const int32_t *rflag;
    const uint32_t *sl, *sr, *nm;
    const int32_t *tid_of_cid;
    const int32_t *f[10];  // flag tid pos mapq sl sr nm ntid npos tlen
};

struct ReadSrc {
    const uint8_t *names;
    const int64_t *name_off;
    const uint8_t *seq;
    const uint8_t *qual;
    const uint32_t *lens;
    const uint32_t *qual_lens;
    int64_t stride;
};

struct RecF {
    int64_t flag, tid, pos, mapq, sl, sr, nm, ntid, npos, tlen;
};

__device__ __forceinline__ int64_t src_row(const FieldSrc &s, int64_t i) {
    return s.se == 2 ? ((i & 1) ? s.np + (i >> 1) : (i >> 1)) : i;
}

__device__ __forceinline__ RecF load_fields(const FieldSrc &s, int64_t i) {
    RecF r;
    if (s.se == 2) {
        // _emit_bam_from_batch PE (linear_gpu_fm.mojo:826-870)
        PeMate m[2];
        pe_pair(s.np, i >> 1, s.cid, s.pos, s.rflag, s.lens, s.tid_of_cid, s.n_tid, m);
        const bool r2 = (i & 1) != 0;  // selects, not an index: the pair stays in registers
        const PeMate q{r2 ? m[1].flag : m[0].flag, r2 ? m[1].tid : m[0].tid,   r2 ? m[1].pos : m[0].pos,
                       r2 ? m[1].ntid : m[0].ntid, r2 ? m[1].npos : m[0].npos, r2 ? m[1].tlen : m[0].tlen};
        const int64_t row = src_row(s, i);
        r.flag = q.flag;
        r.tid = q.tid;
        r.pos = q.pos;
        r.mapq = s.mapq[row];
        r.sl = s.sl[row];
        r.sr = s.sr[row];
        r.nm = s.nm[row];
        r.ntid = q.ntid;
        r.npos = q.npos;
        r.tlen = q.tlen;
    } else if (s.se) {
        // _emit_bam_from_batch SE: flag | 1 | 64; unknown contig -> tid -1, flag | 4.
        const int32_t c = s.cid[i];
        r.flag = (int64_t)s.rflag[i] | 1 | 64;
        r.tid = -1;
        if (c >= 0 && c < s.n_tid)
            r.tid = s.tid_of_cid[c];
        else
            r.flag |= 4;
        r.pos = s.pos[i];
        r.mapq = s.mapq[i];
        r.sl = s.sl[i];
        r.sr = s.sr[i];
        r.nm = s.nm[i];
        r.ntid = -1;
        r.npos = -1;
        r.tlen = 0;
    } else {
        r.flag = s.f[0][i];
        r.tid = s.f[1][i];
        r.pos = s.f[2][i];
        r.mapq = s.f[3][i];
        r.sl = s.f[4][i];
        r.sr = s.f[5][i];
        r.nm = s.f[6][i];
        r.ntid = s.f[7] ? s.f[7][i] : -1;
        r.npos = s.f[8] ? s.f[8][i] : -1;
        r.tlen = s.f[9] ? s.f[9][i] : 0;
    }
    return r;
}

__host__ __device__ __forceinline__ int64_t reg2bin(int64_t beg, int64_t end) {
    if (end <= beg) end = beg + 1;
    end -= 1;
    if ((beg >> 14) == (end >> 14)) return 4681 + (beg >> 14);
    if ((beg >> 17) == (end >> 17)) return 585 + (beg >> 17);
    if ((beg >> 20) == (end >> 20)) return 73 + (beg >> 20);
    if ((beg >> 23) == (end >> 23)) return 9 + (beg >> 23);
    if ((beg >> 26) == (end >> 26)) return 1 + (beg >> 26);
    return 0;
}

// A C G T (either case) -> 1 2 4 8, else 15 (the 4-bit BAM codes of _nt16).
__device__ __forceinline__ uint32_t nt16(uint32_t b) {
    const uint32_t u = b & 0xDF;  // upper case
    return u == 'A' ? 1u : u == 'C' ? 2u : u == 'G' ? 4u : u == 'T' ? 8u : 15u;
}

// 1<->8, 2<->4, else 15 (_nt16_comp): a 4-bit reverse of a one-hot code.
__device__ __forceinline__ uint32_t nt16_comp(uint32_t c) {
    return (c == 1 || c == 2 || c == 4 || c == 8) ? (((c & 1) << 3) | ((c & 2) << 1) | ((c & 4) >> 1) | (c >> 3))
                                                  : 15u;
}

// Byte layout of one record (offsets from the record's block_size field).
struct RecLay {
    uint32_t hdr[9];  // block_size, refID, pos, bin_mq_nl, flag_nc, l_seq, next_refID, next_pos, tlen
    uint32_t cig[3];
    int32_t name_len, qlen, o_cig, o_seq, o_qual, o_tags, total;
    int32_t rev, has_qual, nm8;
};

__device__ __forceinline__ RecLay lay_out(const RecF &f, int64_t name_len, int64_t qlen,
                                          int64_t qual_len, bool qual_ptr, int rg_n) {
    RecLay L;
    const bool rev = (f.flag & 16) != 0;
    int64_t slv = f.sl < 0 ? 0 : f.sl, srv = f.sr < 0 ? 0 : f.sr;
    if (slv + srv >= qlen) slv = srv = 0;
    const int64_t mid = qlen - slv - srv;
    if (rev) {
        const int64_t t = slv;
        slv = srv;
        srv = t;
    }
    const bool mapped = f.tid >= 0 && (f.flag & 4) == 0;
    // CIGAR slS mM srS, the empty operations left out (selects, not a
    // dynamically indexed array, which would live in scratch)
    const bool a = mapped && slv > 0, b = mapped && mid > 0, c = mapped && srv > 0;
    const uint32_t opA = (uint32_t)((slv << 4) | 4), opB = (uint32_t)(mid << 4), opC = (uint32_t)((srv << 4) | 4);
    const uint32_t tailBC = b ? opB : (c ? opC : 0u), tailC = c ? opC : 0u;
    L.cig[0] = a ? opA : tailBC;
    L.cig[1] = a ? tailBC : (b ? tailC : 0u);
    L.cig[2] = (a && b) ? tailC : 0u;
    const int nc = (int)a + (int)b + (int)c;
    const int64_t l_qname = name_len + 1;
    const int64_t block = 32 + l_qname + 4 * nc + (qlen + 1) / 2 + qlen + (3 + rg_n + 1 + 4);
    int64_t bin = 4680, tid_w = -1, pos_w = -1;
    if (mapped) {
        tid_w = f.tid;
        pos_w = f.pos;
        int64_t end = f.pos + mid;
        if (end <= f.pos) end = f.pos + 1;
        bin = reg2bin(f.pos, end);
    }
    int64_t ntid = f.ntid, npos = f.npos;
    if (ntid < 0) ntid = npos = -1;
    const int64_t mq = f.mapq < 0 ? 0 : (f.mapq > 255 ? 255 : f.mapq);
    L.hdr[0] = (uint32_t)block;
    L.hdr[1] = (uint32_t)tid_w;
    L.hdr[2] = (uint32_t)pos_w;
    L.hdr[3] = (uint32_t)((bin << 16) | (mq << 8) | l_qname);
    L.hdr[4] = (uint32_t)((f.flag << 16) | nc);
    L.hdr[5] = (uint32_t)qlen;
    L.hdr[6] = (uint32_t)ntid;
    L.hdr[7] = (uint32_t)npos;
    L.hdr[8] = (uint32_t)f.tlen;
    L.name_len = (int32_t)name_len;
    L.qlen = (int32_t)qlen;
    L.o_cig = (int32_t)(36 + l_qname);
    L.o_seq = L.o_cig + 4 * nc;
    L.o_qual = L.o_seq + (int32_t)((qlen + 1) / 2);
    L.o_tags = L.o_qual + (int32_t)qlen;
    L.total = (int32_t)(4 + block);
    L.rev = rev;
    L.has_qual = qual_ptr && qual_len == qlen;
    L.nm8 = (int32_t)(f.nm < 0 ? 0 : (f.nm > 255 ? 255 : f.nm));
    return L;
}

__device__ __forceinline__ uint32_t word_byte(uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; }

// Byte t of a record's tags: RG:Z<id>\0 then NM:C<nm>.
__device__ __forceinline__ uint32_t tag_byte(int t, const uint8_t *__restrict__ rg, int rg_n, uint32_t nm8) {
    if (t < 3) return t == 0 ? 'R' : (t == 1 ? 'G' : 'Z');
    if (t < 3 + rg_n) return rg[t - 3];
    if (t == 3 + rg_n) return 0u;
    const int u = t - 4 - rg_n;
    return u == 0 ? 'N' : (u == 1 ? 'M' : (u == 2 ? 'C' : nm8));
}

// Byte b of a record (0 <= b < L.total); `words` = the record's 9 header words
// and 3 CIGAR words (LDS: a lane-indexed array would live in scratch).
__device__ __forceinline__ uint32_t rec_byte(const RecLay &L, const uint32_t *words, int b,
                                             const uint8_t *__restrict__ name,
                                             const uint8_t *__restrict__ seq,
                                             const uint8_t *__restrict__ qual,
                                             const uint8_t *__restrict__ rg, int rg_n) {
    if (b < 36) return word_byte(words[b >> 2], b & 3);
    if (b < L.o_cig) {
        const int t = b - 36;
        return t < L.name_len ? (uint32_t)name[t] : 0u;
    }
    if (b < L.o_seq) {
        const int t = b - L.o_cig;
        return word_byte(words[9 + (t >> 2)], t & 3);
    }
    if (b < L.o_qual) {
        const int i = 2 * (b - L.o_seq);
        uint32_t a, c = 0;
        if (L.rev) {
            a = nt16_comp(nt16(seq[L.qlen - 1 - i]));
            if (i + 1 < L.qlen) c = nt16_comp(nt16(seq[L.qlen - 2 - i]));
        } else {
            a = nt16(seq[i]);
            if (i + 1 < L.qlen) c = nt16(seq[i + 1]);
        }
        return (a << 4) | c;
    }
    if (b < L.o_tags) {
        if (!L.has_qual) return 255u;
        const int t = b - L.o_qual;
        const uint32_t q = qual[L.rev ? L.qlen - 1 - t : t];
        return q >= 33 ? q - 33 : 255u;
    }
    return tag_byte(b - L.o_tags, rg, rg_n, (uint32_t)L.nm8);
}

// A wave-uniform 64-bit value into scalar registers (both halves through uint32_t).
__device__ __forceinline__ int64_t uni64(uint64_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return (int64_t)(((uint64_t)hi << 32) | (uint64_t)lo);
}

// One record's layout, 64 bytes, in output order.
struct RecLayout {
    uint32_t w[12];   // header words 0..8 (word 5 = l_seq), CIGAR words 9..11
    uint64_t name;    // first name byte in the names array
    uint32_t row;     // read index (bases and qualities at row * stride)
    uint32_t flags;   // name_len | rev << 16 | has_qual << 17 | n_cigar << 18 | nm8 << 20
};
static_assert(sizeof(RecLayout) == 64, "one 64-byte line per record");

// inv[perm[j]] = j: where each source record goes (entries outside [0, n) are
// skipped; inv starts all ones, so a source that perm never names is skipped too).
__global__ void bam_inverse_kernel(int64_t n, const uint32_t *__restrict__ perm, uint32_t *__restrict__ inv) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t i = perm[j];
        if (i < n) inv[i] = (uint32_t)j;
    }
}

// In source order (the field arrays read in sequence): each record's layout,
// 0x400 applied, written to its output slot.
__global__ void bam_layout_kernel(int64_t n, FieldSrc fs, ReadSrc rs, int rg_n, const uint32_t *__restrict__ inv,
                                  const uint32_t *__restrict__ is_dup, int apply_dup,
                                  RecLayout *__restrict__ lay) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        RecF f = load_fields(fs, i);
        if (apply_dup && is_dup[i] != 0) f.flag |= 1024;
        const int64_t row = src_row(fs, i);
        const int64_t qlen = rs.lens[row];
        const int64_t ql = rs.qual_lens ? rs.qual_lens[row] : qlen;
        const int64_t nb = rs.name_off[row];
        const RecLay L = lay_out(f, rs.name_off[row + 1] - nb, qlen, ql, rs.qual != nullptr, rg_n);
        RecLayout o;
#pragma unroll
        for (int k = 0; k < 9; k++) o.w[k] = L.hdr[k];
#pragma unroll
        for (int k = 0; k < 3; k++) o.w[9 + k] = L.cig[k];
        o.name = (uint64_t)nb;
        o.row = (uint32_t)row;
        o.flags = (uint32_t)L.name_len | (uint32_t)L.rev << 16 | (uint32_t)L.has_qual << 17 |
                  (uint32_t)((L.o_seq - L.o_cig) / 4) << 18 | (uint32_t)L.nm8 << 20;
        const uint32_t j = inv ? inv[i] : (uint32_t)i;
        if (j < n) lay[j] = o;
    }
}

// Record lengths in output order (for the offset scan).
__global__ void bam_sizes_kernel(int64_t n, const RecLayout *__restrict__ lay, uint32_t *__restrict__ sizes) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        sizes[j] = lay[j].w[0] + 4;
}

// The read-group id bytes, from the kernel arguments into device memory (one
// wave), so the pack kernel reads them with ordinary loads.
__global__ void bam_rg_kernel(RgArg rg, uint8_t *__restrict__ dst) {
    for (int t = threadIdx.x; t < BAM_RG_MAX; t += blockDim.x) dst[t] = rg.b[t];
}

// One record's bytes at out + base: each lane builds 4-byte-aligned words
// (byte stores only for the two words shared with the neighbouring records).
__device__ __forceinline__ void write_record(const RecLay &L, const uint32_t *wl, int64_t base,
                                             const uint8_t *name, const uint8_t *seq, const uint8_t *qual,
                                             const uint8_t *rg, int rg_n, int lane, uint8_t *out,
                                             uint32_t *out32) {
    const int64_t d0 = base >> 2, d1 = (base + L.total - 1) >> 2;
    for (int64_t d = d0 + lane; d <= d1; d += 64) {
        const int b0 = (int)(4 * d - base);
        if (b0 >= 0 && b0 + 4 <= L.total) {
            uint32_t w = rec_byte(L, wl, b0, name, seq, qual, rg, rg_n);
            w |= rec_byte(L, wl, b0 + 1, name, seq, qual, rg, rg_n) << 8;
            w |= rec_byte(L, wl, b0 + 2, name, seq, qual, rg, rg_n) << 16;
            w |= rec_byte(L, wl, b0 + 3, name, seq, qual, rg, rg_n) << 24;
            out32[d] = w;
        } else {
            for (int k = 0; k < 4; k++) {
                const int b = b0 + k;
                if (b >= 0 && b < L.total) out[base + b] = (uint8_t)rec_byte(L, wl, b, name, seq, qual, rg, rg_n);
            }
        }
    }
}

// LDS writes of the wave visible to its own later LDS reads.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A record's image: its bytes at the output's alignment (image byte a + i is
// record byte i, a = the record's output offset mod 4).
constexpr int IMG_BYTES = 4 + 36 + (STAGE + 1) + 12 + (STAGE + 1) / 2 + STAGE + 3 + BAM_RG_MAX + 1 + 4;
constexpr int IMG_WORDS = (IMG_BYTES + 7) / 8;

__global__ __launch_bounds__(256) void bam_pack_kernel(int64_t n, ReadSrc rs, const RecLayout *__restrict__ lay,
                                                       const uint8_t *__restrict__ rg, int rg_n,
                                                       const uint64_t *__restrict__ rec_off,
                                                       uint8_t *__restrict__ out, int64_t out_cap) {
    // Each wave takes 64 consecutive output records: their layouts (one 64-byte
    // line each, read in sequence) go to LDS, then the wave writes the 64
    // records one after another. A record is assembled in LDS section by
    // section (header, name, CIGAR, bases from 4-bit codes, qualities, tags; a
    // byte per lane), then leaves as aligned 4-byte words, byte stores only for
    // the words it shares with its neighbours. Its name, bases and qualities
    // load one record ahead, into registers, while the record before it is
    // written.
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t wave = (int64_t)__builtin_amdgcn_readfirstlane(blockIdx.x * BAM_WPB + wv);
    const int64_t n_waves = (int64_t)gridDim.x * BAM_WPB;
    uint32_t *const out32 = reinterpret_cast<uint32_t *>(out);
    __shared__ RecLayout s_rec[BAM_WPB][64];
    __shared__ uint64_t s_base[BAM_WPB][64];
    __shared__ uint64_t s_img[BAM_WPB][IMG_WORDS];
    __shared__ uint8_t s_code[BAM_WPB][STAGE + 8];
    RecLayout *const slots = s_rec[wv];
    uint64_t *const bases = s_base[wv];
    uint8_t *const img = reinterpret_cast<uint8_t *>(s_img[wv]);
    uint8_t *const code = s_code[wv];
    for (int64_t j0 = wave * 64; j0 < n; j0 += n_waves * 64) {
        const int64_t j = j0 + lane;
        if (j < n) {
            slots[lane] = lay[j];
            bases[lane] = rec_off[j];
        }
        wsync();
        const int cnt = (int)(n - j0 < 64 ? n - j0 : 64);
        // record r's name, bases and qualities: bytes lane + 64 k of each
        uint32_t pn[STAGE / 64], ps[STAGE / 64], pq[STAGE / 64];
        auto fetch = [&](int r) {
            const RecLayout &S = slots[r];
            const uint32_t fl = __builtin_amdgcn_readfirstlane(S.flags);
            const int nl = (int)(fl & 0xFFFF), ql = (int)__builtin_amdgcn_readfirstlane(S.w[5]);
            const int64_t row = (int64_t)__builtin_amdgcn_readfirstlane(S.row);
            const uint8_t *nm = rs.names + uni64(S.name);
            const uint8_t *sq = rs.seq + row * rs.stride;
            const uint8_t *qu = ((fl >> 17) & 1) ? rs.qual + row * rs.stride : nullptr;
#pragma unroll
            for (int k = 0; k < STAGE / 64; k++) {
                const int x = lane + 64 * k;
                pn[k] = x < nl ? nm[x] : 0u;
                ps[k] = x < ql ? sq[x] : 0u;
                pq[k] = (qu && x < ql) ? qu[x] : 0u;
            }
        };
        if (cnt > 0) fetch(0);
        for (int r = 0; r < cnt; r++) {
            const RecLayout &S = slots[r];
            const uint32_t fl = __builtin_amdgcn_readfirstlane(S.flags);
            RecLay L;
            L.qlen = (int32_t)__builtin_amdgcn_readfirstlane(S.w[5]);
            L.rev = (fl >> 16) & 1;
            L.has_qual = (fl >> 17) & 1;
            L.name_len = (int32_t)(fl & 0xFFFF);
            L.nm8 = (int32_t)((fl >> 20) & 0xFF);
            L.o_cig = 36 + L.name_len + 1;
            L.o_seq = L.o_cig + 4 * (int32_t)((fl >> 18) & 3);
            L.o_qual = L.o_seq + (L.qlen + 1) / 2;
            L.o_tags = L.o_qual + L.qlen;
            L.total = (int32_t)__builtin_amdgcn_readfirstlane(S.w[0]) + 4;
            const int64_t base = uni64(bases[r]);
            const int64_t row = (int64_t)__builtin_amdgcn_readfirstlane(S.row);
            const bool fits = L.name_len <= STAGE && L.qlen <= STAGE;
            const bool room = base + L.total <= out_cap;  // else the caller sees rec_off[n] > out_cap
            if (fits && room) {
                const int a = (int)(base & 3);
                uint8_t *const im = img + a;
                if (lane < 36) im[lane] = (uint8_t)word_byte(S.w[lane >> 2], lane & 3);
                const int nc4 = L.o_seq - L.o_cig;
                if (lane < nc4) im[L.o_cig + lane] = (uint8_t)word_byte(S.w[9 + (lane >> 2)], lane & 3);
                if (lane == 0) im[36 + L.name_len] = 0;
#pragma unroll
                for (int k = 0; k < STAGE / 64; k++) {
                    const int x = lane + 64 * k;
                    if (x < L.name_len) im[36 + x] = (uint8_t)pn[k];
                    if (x < L.qlen) {
                        // output base / quality index of source byte x
                        const int pos = L.rev ? L.qlen - 1 - x : x;
                        const uint32_t c = nt16(ps[k]);
                        code[pos] = (uint8_t)(L.rev ? nt16_comp(c) : c);
                        const uint32_t q = pq[k];
                        im[L.o_qual + pos] = (uint8_t)(L.has_qual ? (q >= 33 ? q - 33 : 255u) : 255u);
                    }
                }
                if (lane == 0) code[L.qlen] = 0;  // the low half of an odd length's last byte
                const int n_tag = 3 + rg_n + 1 + 4;
                for (int t = lane; t < n_tag; t += 64) im[L.o_tags + t] = (uint8_t)tag_byte(t, rg, rg_n, (uint32_t)L.nm8);
                wsync();
                if (r + 1 < cnt) fetch(r + 1);  // in flight while this record leaves
                const int n_sb = (L.qlen + 1) / 2;
                for (int t = lane; t < n_sb; t += 64)
                    im[L.o_seq + t] = (uint8_t)((code[2 * t] << 4) | code[2 * t + 1]);
                wsync();
                // out: word d holds record bytes 4d - base .. + 3, image bytes 4 (d - d0) ..
                const int64_t d0 = base >> 2, d1 = (base + L.total - 1) >> 2;
                for (int64_t d = d0 + lane; d <= d1; d += 64) {
                    const int i0 = (int)(4 * d - base);
                    const int m = (int)(4 * (d - d0));
                    if (i0 >= 0 && i0 + 4 <= L.total) {
                        out32[d] = *reinterpret_cast<const uint32_t *>(img + m);
                    } else {
                        for (int k = 0; k < 4; k++) {
                            const int b = i0 + k;
                            if (b >= 0 && b < L.total) out[base + b] = img[m + k];
                        }
                    }
                }
            } else {
                if (r + 1 < cnt) fetch(r + 1);
                // longer than the stage: every byte from HBM
                if (room)
                    write_record(L, S.w, base, rs.names + uni64(S.name), rs.seq + row * rs.stride,
                                 rs.qual ? rs.qual + row * rs.stride : nullptr, rg, rg_n, lane, out, out32);
            }
            wsync();  // the image and codes are rewritten for the next record
        }
        wsync();  // the slots are reused by the next batch
    }
}

struct BamScratch {
    size_t sizes, partial, rg, inv, lay, total;
};

BamScratch bam_layout(int64_t n) {
    BamScratch s;
    size_t off = 0;
    auto take = [&](size_t b) {
        size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    s.sizes = take((size_t)(n < 1 ? 1 : n) * 4);
    s.partial = take(scan_scratch_bytes(n < 1 ? 1 : n));
    s.rg = take(BAM_RG_MAX);
    s.inv = take((size_t)(n < 1 ? 1 : n) * 4);
    s.lay = take((size_t)(n < 1 ? 1 : n) * sizeof(RecLayout));
    s.total = off;
    return s;
}

int bam_pack(void *stream, int64_t n, const FieldSrc &fs, const ma_bam_reads *r, const char *rg,
             int32_t rg_n, const uint32_t *d_perm, const uint32_t *d_is_dup, int32_t apply_dup,
             uint8_t *d_out, int64_t out_cap, uint64_t *d_rec_off, void *d_scratch,
             int64_t scratch_bytes, const char *who) {
    MA_REQUIRE(n >= 0 && r && d_rec_off, std::string(who) + ": bad arguments");
    MA_REQUIRE(n <= 0xFFFFFFFFll, std::string(who) + ": n exceeds u32 record ids");
    MA_REQUIRE(rg_n >= 0 && rg_n <= BAM_RG_MAX && (rg_n == 0 || rg),
               std::string(who) + ": read-group id longer than 240 bytes");
    MA_REQUIRE(!apply_dup || d_is_dup, std::string(who) + ": apply_dup without is_dup");
    MA_REQUIRE(((uintptr_t)d_out & 3) == 0, std::string(who) + ": output not 4-byte aligned");
    BamScratch S = bam_layout(n);
    MA_REQUIRE(scratch_bytes >= (int64_t)S.total, std::string(who) + ": scratch too small");
    hipStream_t st = (hipStream_t)stream;
    if (n == 0) return fill_u64(st, d_rec_off, 1, 0);
    MA_REQUIRE(r->names && r->name_off && r->seq && r->lens, std::string(who) + ": null read array");
    ReadSrc rs{r->names, r->name_off, r->seq, r->qual, r->lens, r->qual_lens, r->stride};
    RgArg ra;
    ra.n = rg_n;
    for (int k = 0; k < BAM_RG_MAX; k++) ra.b[k] = k < rg_n ? (uint8_t)rg[k] : 0;
    char *s = (char *)d_scratch;
    uint32_t *sizes = (uint32_t *)(s + S.sizes);
    uint32_t *inv = d_perm ? (uint32_t *)(s + S.inv) : nullptr;
    RecLayout *lay = (RecLayout *)(s + S.lay);
    if (inv) {
        // a permutation fills every slot; anything else leaves empty layouts (4-byte records), never a wild read
        MA_TRY(fill_u64(st, (uint64_t *)inv, (n + 1) / 2, ~0ull));
        MA_TRY(fill_u64(st, (uint64_t *)lay, n * (int64_t)(sizeof(RecLayout) / 8), 0));
        bam_inverse_kernel<<<grid_for(n), 256, 0, st>>>(n, d_perm, inv);
        MA_LAUNCHED();
    }
    bam_layout_kernel<<<grid_for(n), 256, 0, st>>>(n, fs, rs, rg_n, inv, d_is_dup, apply_dup, lay);
    MA_LAUNCHED();
    bam_sizes_kernel<<<grid_for(n), 256, 0, st>>>(n, lay, sizes);
    MA_LAUNCHED();
    MA_TRY(exclusive_scan_u32(st, n, sizes, d_rec_off, (uint64_t *)(s + S.partial), d_rec_off + n));
    uint8_t *d_rg = (uint8_t *)(s + S.rg);
    bam_rg_kernel<<<1, 64, 0, st>>>(ra, d_rg);
    MA_LAUNCHED();
    const int64_t blocks = std::min<int64_t>((n + 64 * BAM_WPB - 1) / (64 * BAM_WPB), 256 * 16);
    bam_pack_kernel<<<(unsigned)blocks, 256, 0, st>>>(n, rs, lay, d_rg, rg_n, d_rec_off, d_out, out_cap);
    MA_LAUNCHED();
    return MA_OK;
}

// ------------------------------------------------------------------ BGZF (host)
constexpr uint8_t BGZF_EOF[28] = {0x1f, 0x8b, 0x08, 0x04, 0x00, 0x00, 0x00, 0x00, 0x00, 0xff,
                                  0x06, 0x00, 0x42, 0x43, 0x02, 0x00, 0x1b, 0x00, 0x03, 0x00,
                                  0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00};
constexpr int64_t BGZF_MEMBER_MAX = 65536;

void put_le16(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
}
void put_le32(uint8_t *p, uint32_t v) {
    for (int k = 0; k < 4; k++) p[k] = (uint8_t)(v >> (8 * k));
}

// One member (header + raw deflate + CRC32 + ISIZE) of src[0..n) into dst
// (BGZF_MEMBER_MAX bytes); returns its length or -1.
int64_t bgzf_member(const uint8_t *src, int64_t n, int level, uint8_t *dst) {
    z_stream zs{};
    if (deflateInit2(&zs, level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) return -1;
    // The reference compresses with Python's zlib object: compress() (Z_NO_FLUSH)
    // then flush() (Z_FINISH), each into an output buffer that starts at 32 KiB
    // and grows by 64 KiB while deflate fills it. Stored blocks (level 0) are cut
    // at those buffer ends, so the same call pattern gives the same bytes.
    zs.next_in = const_cast<Bytef *>(src);
    zs.avail_in = (uInt)n;
    uint8_t *const out0 = dst + 18;
    const int64_t room = BGZF_MEMBER_MAX - 18 - 8;
    int rc = Z_OK;
    for (int phase = 0; phase < 2 && rc != Z_STREAM_ERROR; phase++) {
        int64_t grant = 32768;
        do {
            const int64_t used = (int64_t)zs.total_out;
            const int64_t give = std::min(grant, room - used);
            if (give <= 0) {
                rc = Z_STREAM_ERROR;  // the member would pass 64 KiB
                break;
            }
            zs.next_out = out0 + used;
            zs.avail_out = (uInt)give;
            rc = deflate(&zs, phase == 0 ? Z_NO_FLUSH : Z_FINISH);
            grant = 65536;
        } while (zs.avail_out == 0 && rc != Z_STREAM_ERROR && rc != Z_STREAM_END);
    }
    const int64_t clen = (int64_t)zs.total_out;
    deflateEnd(&zs);
    if (rc != Z_STREAM_END) return -1;
    static const uint8_t head[16] = {31, 139, 8, 4, 0, 0, 0, 0, 0, 255, 6, 0, 66, 67, 2, 0};
    for (int k = 0; k < 16; k++) dst[k] = head[k];
    const int64_t total = 18 + clen + 8;
    put_le16(dst + 16, (uint32_t)(total - 1));
    put_le32(dst + 18 + clen, (uint32_t)crc32(0L, src, (uInt)n));
    put_le32(dst + 18 + clen + 4, (uint32_t)n);
    return total;
}

// Compresses src into consecutive 65280-byte members with up to `threads`
// workers; worker t's members, in order, end up in parts[t] (so parts[0],
// parts[1], ... concatenated is the BGZF stream). Returns false on a zlib error.
bool bgzf_compress_parts(const uint8_t *src, int64_t n, int level, int threads,
                         std::vector<std::vector<uint8_t>> &parts) {
    const int64_t nb = (n + BGZF_BLOCK - 1) / BGZF_BLOCK;
    parts.clear();
    if (nb == 0) return true;
    const int T = std::max(1, std::min<int>(threads, (int)std::min<int64_t>(nb, 256)));
    parts.resize(T);
    std::vector<int> ok(T, 1);
    auto work = [&](int t) {
        const int64_t b0 = nb * t / T, b1 = nb * (t + 1) / T;
        parts[t].resize((size_t)(b1 - b0) * BGZF_MEMBER_MAX);
        int64_t off = 0;
        for (int64_t b = b0; b < b1; b++) {
            const int64_t s = b * BGZF_BLOCK, len = std::min(BGZF_BLOCK, n - s);
            const int64_t m = bgzf_member(src + s, len, level, parts[t].data() + off);
            if (m < 0) {
                ok[t] = 0;
                return;
            }
            off += m;
        }
        parts[t].resize((size_t)off);
    };
    if (T == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < T; t++) pool.emplace_back(work, t);
        for (auto &th : pool) th.join();
    }
    for (int t = 0; t < T; t++)
        if (!ok[t]) return false;
    return true;
}

}  // namespace

struct ma_bgzf {
    FILE *fh = nullptr;
    int level = 1, threads = 1;
    std::vector<uint8_t> tail;  // < 65280 bytes waiting for the next member
    std::vector<std::vector<uint8_t>> parts;  // compressed members of one write, per worker
    int64_t written = 0;
};

extern "C" {

int64_t ma_bam_scratch_bytes(int64_t n) { return (int64_t)bam_layout(n).total; }

int64_t ma_bam_record_bound(int32_t name_len, int32_t len, int32_t rg_n) {
    return 36 + (int64_t)name_len + 1 + 12 + ((int64_t)len + 1) / 2 + len + 3 + rg_n + 1 + 4;
}

int ma_bam_pack_dev(void *stream, int64_t n, const ma_bam_fields *f, const ma_bam_reads *r,
                    const char *rg, int32_t rg_n, const uint32_t *d_perm, const uint32_t *d_is_dup,
                    int32_t apply_dup, uint8_t *d_out, int64_t out_cap, uint64_t *d_rec_off,
                    void *d_scratch, int64_t scratch_bytes) {
    MA_REQUIRE(f, "ma_bam_pack_dev: null fields");
    MA_REQUIRE(n == 0 || (f->flag && f->tid && f->pos && f->mapq && f->sl && f->sr && f->nm),
               "ma_bam_pack_dev: null field array");
    FieldSrc fs{};
    fs.se = 0;
    const int32_t *arr[10] = {f->flag, f->tid, f->pos, f->mapq, f->sl, f->sr, f->nm, f->ntid, f->npos, f->tlen};
    for (int k = 0; k < 10; k++) fs.f[k] = arr[k];
    return bam_pack(stream, n, fs, r, rg, rg_n, d_perm, d_is_dup, apply_dup, d_out, out_cap, d_rec_off,
                    d_scratch, scratch_bytes, "ma_bam_pack_dev");
}

int ma_bam_pack_se_dev(void *stream, int64_t n, const ma_recs *m, const int32_t *d_tid_of_cid,
                       int32_t n_tid, const ma_bam_reads *r, const char *rg, int32_t rg_n,
                       const uint32_t *d_perm, const uint32_t *d_is_dup, int32_t apply_dup,
                       uint8_t *d_out, int64_t out_cap, uint64_t *d_rec_off, void *d_scratch,
                       int64_t scratch_bytes) {
    MA_REQUIRE(m && (n_tid == 0 || d_tid_of_cid) && n_tid >= 0, "ma_bam_pack_se_dev: bad arguments");
    FieldSrc fs{};
    fs.se = 1;
    fs.n_tid = n_tid;
    fs.cid = m->cid;
    fs.pos = m->pos;
    fs.mapq = m->mapq;
    fs.rflag = m->flag;
    fs.sl = m->sl;
    fs.sr = m->sr;
    fs.nm = m->nm;
    fs.tid_of_cid = d_tid_of_cid;
    return bam_pack(stream, n, fs, r, rg, rg_n, d_perm, d_is_dup, apply_dup, d_out, out_cap, d_rec_off,
                    d_scratch, scratch_bytes, "ma_bam_pack_se_dev");
}

int ma_bam_pack_pe_dev(void *stream, int64_t n_pairs, const ma_recs *m, const int32_t *d_tid_of_cid,
                       int32_t n_tid, const ma_bam_reads *r, const char *rg, int32_t rg_n,
                       const uint32_t *d_perm, const uint32_t *d_is_dup, int32_t apply_dup,
                       uint8_t *d_out, int64_t out_cap, uint64_t *d_rec_off, void *d_scratch,
                       int64_t scratch_bytes) {
    MA_REQUIRE(m && r && n_pairs >= 0 && (n_tid == 0 || d_tid_of_cid) && n_tid >= 0,
               "ma_bam_pack_pe_dev: bad arguments");
    MA_REQUIRE(n_pairs == 0 || r->lens, "ma_bam_pack_pe_dev: null read lengths");
    FieldSrc fs{};
    fs.se = 2;
    fs.n_tid = n_tid;
    fs.np = n_pairs;
    fs.lens = r->lens;
    fs.cid = m->cid;
    fs.pos = m->pos;
    fs.mapq = m->mapq;
    fs.rflag = m->flag;
    fs.sl = m->sl;
    fs.sr = m->sr;
    fs.nm = m->nm;
    fs.tid_of_cid = d_tid_of_cid;
    return bam_pack(stream, 2 * n_pairs, fs, r, rg, rg_n, d_perm, d_is_dup, apply_dup, d_out, out_cap, d_rec_off,
                    d_scratch, scratch_bytes, "ma_bam_pack_pe_dev");
}

int64_t ma_bgzf_bound(int64_t n) {
    const int64_t nb = (n + BGZF_BLOCK - 1) / BGZF_BLOCK;
    return nb * BGZF_MEMBER_MAX + 28;
}

int ma_bgzf_compress(const uint8_t *src, int64_t n, int32_t level, int32_t threads, int32_t add_eof,
                     uint8_t *dst, int64_t dst_cap, int64_t *out_n) {
    MA_REQUIRE(n >= 0 && (n == 0 || src) && dst && out_n, "ma_bgzf_compress: bad arguments");
    MA_REQUIRE(level >= 0 && level <= 9, "ma_bgzf_compress: level outside 0..9");
    std::vector<std::vector<uint8_t>> parts;
    MA_REQUIRE(bgzf_compress_parts(src, n, level, threads, parts), "ma_bgzf_compress: deflate failed");
    int64_t total = 0;
    for (auto &p : parts) total += (int64_t)p.size();
    MA_REQUIRE(total + (add_eof ? 28 : 0) <= dst_cap, "ma_bgzf_compress: output buffer too small");
    total = 0;
    for (auto &p : parts) {
        std::copy(p.begin(), p.end(), dst + total);
        total += (int64_t)p.size();
    }
    if (add_eof) {
        std::copy(BGZF_EOF, BGZF_EOF + 28, dst + total);
        total += 28;
    }
    *out_n = total;
    return MA_OK;
}

int ma_bgzf_inflate(const uint8_t *src, int64_t n, int32_t threads, uint8_t *dst, int64_t dst_cap,
                    int64_t *used, int64_t *out_n) {
    MA_REQUIRE(n >= 0 && (n == 0 || src) && used && out_n, "ma_bgzf_inflate: bad arguments");
    // the whole members at the front of src: a gzip header whose extra field
    // holds BC (BSIZE = member bytes - 1), raw deflate, CRC32, ISIZE
    std::vector<int64_t> soff{0}, doff{0};
    int64_t at = 0;
    while (n - at >= 18) {
        const uint8_t *m = src + at;
        if (m[0] != 31 || m[1] != 139 || m[2] != 8 || m[3] != 4) {
            set_error("ma_bgzf_inflate: not a BGZF member at byte " + std::to_string(at));
            return MA_ERR_ARG;
        }
        const int64_t xlen = m[10] | (int64_t)m[11] << 8;
        if (n - at < 12 + xlen) break;
        int64_t bsize = -1;
        for (int64_t x = 12; x + 4 <= 12 + xlen;) {
            const int64_t slen = m[x + 2] | (int64_t)m[x + 3] << 8;
            if (m[x] == 66 && m[x + 1] == 67 && slen == 2 && x + 6 <= 12 + xlen) bsize = (m[x + 4] | (int64_t)m[x + 5] << 8) + 1;
            x += 4 + slen;
        }
        if (bsize < 12 + xlen + 8) {
            set_error("ma_bgzf_inflate: member without a BGZF block size at byte " + std::to_string(at));
            return MA_ERR_ARG;
        }
        if (n - at < bsize) break;
        const uint8_t *t = m + bsize - 4;
        const int64_t isize = t[0] | (int64_t)t[1] << 8 | (int64_t)t[2] << 16 | (int64_t)t[3] << 24;
        if (isize > 65536) {  // a BGZF member holds at most 64 KiB
            set_error("ma_bgzf_inflate: corrupt BGZF member at byte " + std::to_string(at) + " (ISIZE)");
            return MA_ERR_IO;
        }
        at += bsize;
        soff.push_back(at);
        doff.push_back(doff.back() + isize);
    }
    *used = at;
    *out_n = doff.back();
    if (!dst) return MA_OK;
    MA_REQUIRE(dst_cap >= doff.back(), "ma_bgzf_inflate: output buffer too small");
    const int64_t nm = (int64_t)soff.size() - 1;
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(threads, 1), nm));
    std::vector<int64_t> bad(T, -1);
    auto work = [&](int w) {
        for (int64_t i = nm * w / T; i < nm * (w + 1) / T; i++) {
            const uint8_t *m = src + soff[i];
            const int64_t xlen = m[10] | (int64_t)m[11] << 8;
            const int64_t clen = soff[i + 1] - soff[i] - 12 - xlen - 8;
            const int64_t dlen = doff[i + 1] - doff[i];
            z_stream zs{};
            if (inflateInit2(&zs, -15) != Z_OK) {
                bad[w] = i;
                return;
            }
            zs.next_in = const_cast<Bytef *>(m + 12 + xlen);
            zs.avail_in = (uInt)clen;
            zs.next_out = dst + doff[i];
            zs.avail_out = (uInt)dlen;
            const int rc = inflate(&zs, Z_FINISH);
            const bool ok = rc == Z_STREAM_END && (int64_t)zs.total_out == dlen;
            inflateEnd(&zs);
            const uint8_t *c = src + soff[i + 1] - 8;
            const uint32_t crc = c[0] | (uint32_t)c[1] << 8 | (uint32_t)c[2] << 16 | (uint32_t)c[3] << 24;
            if (!ok || (uint32_t)crc32(0L, dst + doff[i], (uInt)dlen) != crc) {
                bad[w] = i;
                return;
            }
        }
    };
    if (T == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int w = 0; w < T; w++) pool.emplace_back(work, w);
        for (auto &th : pool) th.join();
    }
    for (int w = 0; w < T; w++)
        if (bad[w] >= 0) {
            set_error("ma_bgzf_inflate: corrupt BGZF member at byte " + std::to_string(soff[bad[w]]) +
                      " (inflate, ISIZE or CRC32)");
            return MA_ERR_IO;
        }
    return MA_OK;
}

int ma_gz_open(const char *path, ma_gz **out) {
    MA_REQUIRE(path && out, "ma_gz_open: bad arguments");
    gzFile g = gzopen(path, "rb");
    if (!g) {
        set_error(std::string("ma_gz_open: cannot open ") + path);
        return MA_ERR_IO;
    }
    gzbuffer(g, 1u << 20);
    *out = reinterpret_cast<ma_gz *>(g);
    return MA_OK;
}

int ma_gz_read(ma_gz *f, uint8_t *dst, int64_t cap, int64_t *got) {
    MA_REQUIRE(f && (cap == 0 || dst) && cap >= 0 && got, "ma_gz_read: bad arguments");
    gzFile g = reinterpret_cast<gzFile>(f);
    int64_t n = 0;
    while (n < cap) {
        const int k = gzread(g, dst + n, (unsigned)std::min<int64_t>(cap - n, 1 << 30));
        if (k > 0 && n == 0 && gzdirect(g)) {
            set_error("ma_gz_read: not a gzip file");
            return MA_ERR_ARG;
        }
        if (k <= 0) {
            int err = Z_OK;
            const char *msg = gzerror(g, &err);
            if (k < 0 || (err != Z_OK && err != Z_STREAM_END)) {
                set_error(std::string("ma_gz_read: ") + (msg ? msg : "inflate failed"));
                return MA_ERR_IO;
            }
            break;
        }
        n += k;
    }
    *got = n;
    return MA_OK;
}

int ma_gz_close(ma_gz *f) {
    if (f) gzclose_r(reinterpret_cast<gzFile>(f));
    return MA_OK;
}

int ma_bgzf_open(const char *path, int32_t level, int32_t threads, ma_bgzf **out) {
    MA_REQUIRE(path && out, "ma_bgzf_open: bad arguments");
    FILE *fh = fopen(path, "wb");
    if (!fh) {
        set_error(std::string("ma_bgzf_open: cannot open ") + path);
        return MA_ERR_IO;
    }
    ma_bgzf *w = new ma_bgzf;
    w->fh = fh;
    w->level = std::max(0, std::min(9, (int)level));
    w->threads = std::max(1, (int)threads);
    *out = w;
    return MA_OK;
}

namespace {
int bgzf_emit(ma_bgzf *w, const uint8_t *src, int64_t n) {
    MA_REQUIRE(bgzf_compress_parts(src, n, w->level, w->threads, w->parts), "ma_bgzf_write: deflate failed");
    for (auto &p : w->parts) {
        if (!p.empty() && fwrite(p.data(), 1, p.size(), w->fh) != p.size()) {
            set_error("ma_bgzf_write: write failed");
            return MA_ERR_IO;
        }
        w->written += (int64_t)p.size();
    }
    return MA_OK;
}
}  // namespace

int ma_bgzf_write(ma_bgzf *w, const void *data, int64_t n) {
    MA_REQUIRE(w && w->fh && n >= 0 && (n == 0 || data), "ma_bgzf_write: bad arguments");
    const uint8_t *p = (const uint8_t *)data;
    if (!w->tail.empty()) {
        const int64_t take = std::min<int64_t>(n, BGZF_BLOCK - (int64_t)w->tail.size());
        w->tail.insert(w->tail.end(), p, p + take);
        p += take;
        n -= take;
        if ((int64_t)w->tail.size() < BGZF_BLOCK) return MA_OK;
        MA_TRY(bgzf_emit(w, w->tail.data(), BGZF_BLOCK));
        w->tail.clear();
    }
    const int64_t full = n / BGZF_BLOCK * BGZF_BLOCK;
    if (full) MA_TRY(bgzf_emit(w, p, full));
    w->tail.assign(p + full, p + n);
    return MA_OK;
}

int ma_bgzf_close(ma_bgzf *w, int64_t *bytes_written) {
    MA_REQUIRE(w, "ma_bgzf_close: null writer");
    int rc = MA_OK;
    if (w->fh) {
        if (!w->tail.empty()) rc = bgzf_emit(w, w->tail.data(), (int64_t)w->tail.size());
        if (rc == MA_OK && fwrite(BGZF_EOF, 1, 28, w->fh) != 28) {
            set_error("ma_bgzf_close: write failed");
            rc = MA_ERR_IO;
        }
        w->written += 28;
        if (fclose(w->fh) != 0 && rc == MA_OK) {
            set_error("ma_bgzf_close: close failed");
            rc = MA_ERR_IO;
        }
    }
    if (bytes_written) *bytes_written = w->written;
    delete w;
    return rc;
}

}  // extern "C"
