// FASTQ ingest on the device: raw FASTQ bytes in HBM -> the read batch the
// map and the BAM pack take (rows of original bases padded with 'N', rows of
// converted bases, quality rows, lengths, names back to back).
//
// Semantics follow the reference's bulk reader (fq2bam-meth/src/linear_fastq.mojo):
// lines end at '\n' with one trailing '\r' dropped (_pipe_line, :229-266), an
// unterminated last line counts at the end of the input when non-empty; a
// record is four lines (_store_rec, :342-394; truncation is an error); the
// read name skips a leading '@' and stops at ' ' or '/' (_qname_span, :322-340);
// conversion C/c -> T or G/g -> A (_cvt, :37-46); rows padded with 'N'
// (fq_pack_bases, :432-470).
//
// Design: the buffer is cut into 16 KiB tiles (a 256-thread block, 64 bytes a
// thread): one sweep counts the tile's newlines, a scan gives every tile its
// first line number, a second sweep writes each newline's position at its line
// number; one thread then settles the record count, the bytes consumed and the
// status; a wave per record measures its four lines and name and writes its
// base, converted-base and quality rows (4-byte words gathered from the text,
// converted four bytes at a time, stored whole except where a word is shared
// with the neighbouring row); a scan of the name lengths places the names,
// which a last pass copies.
#include "ma_sort.hpp"

using namespace ma;

namespace {

constexpr int FQ_TB = 256;                   // threads per block
constexpr int FQ_BYTES = 64;                 // bytes per thread per sweep
constexpr int FQ_TILE = FQ_TB * FQ_BYTES;    // 16 KiB

enum : uint32_t { FQ_OK = 0, FQ_NO_SEQ = 1, FQ_NO_PLUS = 2, FQ_NO_QUAL = 3, FQ_TOO_LONG = 4, FQ_NAMES = 5 };

// info words (u64): 0 records, 1 bytes consumed, 2 longest read, 3 status,
// 4 name bytes, 5 newlines, 6 lines, 7 first read longer than stride

__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {  // 0x80 in each zero byte of x
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
}

__device__ __forceinline__ uint32_t count_nl(const uint8_t *buf, int64_t n, int64_t p0) {
    uint32_t c = 0;
    if (p0 + FQ_BYTES <= n && ((p0 | (int64_t)(uintptr_t)buf) & 15) == 0) {
        const uint4 *v = reinterpret_cast<const uint4 *>(buf + p0);
#pragma unroll
        for (int k = 0; k < FQ_BYTES / 16; k++) {
            const uint4 w = v[k];
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t x = ws[q] ^ 0x0A0A0A0Au;  // zero bytes where '\n'
                c += __popc(zero_bytes(x));  // exact: no borrow between bytes
            }
        }
    } else {
        for (int64_t p = p0; p < n && p < p0 + FQ_BYTES; p++) c += buf[p] == '\n';
    }
    return c;
}

__device__ __forceinline__ uint32_t block_sum_256(uint32_t v, uint32_t *sh) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    const uint32_t s = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(FQ_TB) void fq_count_kernel(const uint8_t *__restrict__ buf, int64_t n, int64_t tiles,
                                                         uint32_t *__restrict__ counts) {
    __shared__ uint32_t sh[4];
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t p0 = tile * FQ_TILE + (int64_t)threadIdx.x * FQ_BYTES;
        const uint32_t s = block_sum_256(count_nl(buf, n, p0), sh);
        if (threadIdx.x == 0) counts[tile] = s;
    }
}

// nl[L] = position of the newline that ends line L (for L < cap).
// Newlines a tile stages in LDS before its coalesced write-out (16 KiB of
// FASTQ text holds about 200 lines; a tile with more writes them directly).
constexpr int FQ_STAGE = 2048;

__global__ __launch_bounds__(FQ_TB) void fq_lines_kernel(const uint8_t *__restrict__ buf, int64_t n, int64_t tiles,
                                                         const uint64_t *__restrict__ tile_base,
                                                         uint64_t *__restrict__ nl, int64_t cap) {
    __shared__ uint64_t sh[4];
    __shared__ uint16_t stage[FQ_STAGE];  // the tile's newlines as offsets in the tile, in order
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t t0 = tile * FQ_TILE;
        const int64_t p0 = t0 + (int64_t)threadIdx.x * FQ_BYTES;
        // the thread's 64 bytes: newline flags of each word kept for the writes
        const bool vec = p0 + FQ_BYTES <= n && ((p0 | (int64_t)(uintptr_t)buf) & 15) == 0;
        uint32_t hit[FQ_BYTES / 4];
        uint32_t c = 0;
        if (vec) {
            const uint4 *v = reinterpret_cast<const uint4 *>(buf + p0);
#pragma unroll
            for (int k = 0; k < FQ_BYTES / 16; k++) {
                const uint4 w = v[k];
                const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    hit[4 * k + q] = zero_bytes(ws[q] ^ 0x0A0A0A0Au);  // 0x80 in each '\n' byte
                    c += __popc(hit[4 * k + q]);
                }
            }
        } else {
            c = count_nl(buf, n, p0);
        }
        uint64_t tot;
        const uint64_t ex = block_excl_scan_256(c, sh, tot);
        const uint64_t base = tile_base[tile];
        if (tot <= (uint64_t)FQ_STAGE) {  // block-uniform
            uint32_t L = (uint32_t)ex;
            if (c) {
                if (vec) {
#pragma unroll
                    for (int i = 0; i < FQ_BYTES / 4; i++) {
                        uint32_t h = hit[i];
                        while (h) {
                            stage[L++] = (uint16_t)(threadIdx.x * FQ_BYTES + 4 * i + (__builtin_ctz(h) >> 3));
                            h &= h - 1;
                        }
                    }
                } else {
                    for (int64_t p = p0; p < n && p < p0 + FQ_BYTES; p++)
                        if (buf[p] == '\n') stage[L++] = (uint16_t)(p - t0);
                }
            }
            __syncthreads();
            for (int i = threadIdx.x; i < (int)tot; i += FQ_TB)
                if (base + i < (uint64_t)cap) nl[base + i] = (uint64_t)(t0 + stage[i]);
            // the next tile's scan synchronises the block before stage is written again
        } else {
            uint64_t L = base + ex;
            if (c && L < (uint64_t)cap) {
                if (vec) {
#pragma unroll
                    for (int i = 0; i < FQ_BYTES / 4; i++) {
                        uint32_t h = hit[i];
                        while (h) {
                            const int b = __builtin_ctz(h) >> 3;
                            if (L < (uint64_t)cap) nl[L] = (uint64_t)(p0 + 4 * i + b);
                            L++;
                            h &= h - 1;
                        }
                    }
                } else {
                    for (int64_t p = p0; p < n && p < p0 + FQ_BYTES; p++)
                        if (buf[p] == '\n') {
                            if (L < (uint64_t)cap) nl[L] = (uint64_t)p;
                            L++;
                        }
                }
            }
        }
    }
}

__device__ __forceinline__ int64_t line_start(const uint64_t *nl, int64_t L) { return L == 0 ? 0 : (int64_t)nl[L - 1] + 1; }

// Line L's bytes [s, e) without its '\r'.
__device__ __forceinline__ void line_span(const uint8_t *buf, const uint64_t *nl, int64_t L, int64_t &s, int64_t &e) {
    s = line_start(nl, L);
    e = (int64_t)nl[L];
    if (e > s && buf[e - 1] == '\r') e--;
}

// Record count, bytes consumed and status from the newline count.
__global__ void fq_settle_kernel(const uint8_t *__restrict__ buf, int64_t n, int final, int64_t max_records,
                                 uint64_t *__restrict__ nl, int64_t cap, const uint64_t *__restrict__ total_nl,
                                 uint64_t *__restrict__ info) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint64_t nnl = *total_nl;
    uint64_t lines = nnl;
    if (final && n > 0 && buf[n - 1] != '\n') {
        // an unterminated rest counts as a line when it is not empty after the '\r' strip
        const int64_t s = nnl == 0 ? 0 : (nnl - 1 < (uint64_t)cap ? (int64_t)nl[nnl - 1] + 1 : n);
        int64_t e = n;
        if (e > s && buf[e - 1] == '\r') e--;
        if (e > s && nnl < (uint64_t)cap) {
            nl[nnl] = (uint64_t)n;  // a virtual newline just past the end
            lines = nnl + 1;
        } else if (e > s) {
            lines = nnl + 1;  // beyond the records asked for: never read
        }
    }
    const uint64_t complete = lines / 4;
    const uint64_t recs = complete < (uint64_t)max_records ? complete : (uint64_t)max_records;
    uint32_t status = FQ_OK;
    if (final && recs == complete && lines % 4) status = (uint32_t)(lines % 4);  // 1 no SEQ, 2 no '+', 3 no QUAL
    info[0] = recs;
    uint64_t consumed = recs ? nl[4 * recs - 1] + 1 : 0;
    info[1] = consumed > (uint64_t)n ? (uint64_t)n : consumed;
    info[2] = 0;
    info[3] = status;
    info[4] = 0;
    info[5] = nnl;
    info[6] = lines;
    info[7] = ~0ull;
}

// The 4 text bytes at p (p + 4 <= n): the aligned word holding p and, when p
// is not aligned, the next one (which holds byte p + 3, so both are inside).
__device__ __forceinline__ uint32_t text4(const uint8_t *buf, int64_t p) {
    const uintptr_t a = (uintptr_t)(buf + p);
    const uint32_t *w = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
    const uint32_t s = (uint32_t)(a & 3);
    return s ? __builtin_amdgcn_alignbyte(w[1], w[0], s) : w[0];
}

// Row bytes x0 .. x0 + 3 of a line [ls, ls + len) padded with `pad`: text
// bytes where x < len, pad bytes after.
__device__ __forceinline__ uint32_t row4(const uint8_t *buf, int64_t n, int64_t ls, int64_t len, int64_t x0,
                                         uint32_t pad) {
    const int64_t v = len - x0;  // text bytes among the four
    if (v <= 0) return pad;
    uint32_t w;
    if (ls + x0 >= 0 && ls + x0 + 4 <= n) {
        w = text4(buf, ls + x0);
    } else {
        w = 0;
        for (int j = 0; j < 4; j++) {
            const int64_t q = ls + x0 + j;
            if (q >= 0 && q < n) w |= (uint32_t)buf[q] << (8 * j);
        }
    }
    if (v >= 4) return w;
    const uint32_t keep = (1u << (8 * v)) - 1;
    return (w & keep) | (pad & ~keep);
}

// C/c -> T (mode 1) or G/g -> A (mode 2) in four bytes at once.
__device__ __forceinline__ uint32_t cvt4(uint32_t w, int mode) {
    if (mode == 0) return w;
    const uint32_t up = mode == 1 ? 0x43434343u : 0x47474747u, to = mode == 1 ? 0x54545454u : 0x41414141u;
    const uint32_t hit = zero_bytes(w ^ up) | zero_bytes(w ^ (up | 0x20202020u));
    const uint32_t m = (hit >> 7) * 0xFFu;
    return (w & ~m) | (to & m);
}

// Orders a wave's LDS writes before its later reads by other lanes.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int FQ_CHUNK = 64;  // records per wave step (one per lane while measuring)

struct FqLds {
    int64_t src[4][FQ_CHUNK];  // per wave: the chunk's line starts (bases, then qualities)
    int32_t len[4][FQ_CHUNK];
    int64_t off[4][FQ_CHUNK + 1];  // name offsets
};

// Word v at wa, or the bytes of it that lie in [0, stride) of a row at a0.
__device__ __forceinline__ void put_word(uintptr_t wa, uintptr_t a0, int64_t x0, int64_t stride, uint32_t v) {
    if (x0 >= 0 && x0 + 4 <= stride) {
        *reinterpret_cast<uint32_t *>(wa) = v;
    } else {
        uint8_t *dst = reinterpret_cast<uint8_t *>(a0);
        for (int q = 0; q < 4; q++)
            if (x0 + q >= 0 && x0 + q < stride) dst[x0 + q] = (uint8_t)(v >> (8 * q));
    }
}

// The chunk's rows of one kind, `stride` bytes each from row0 on (converted by
// pmode), and copies converted by cmode from crow0 on when crow0 is not null
// (with row0's alignment): the lanes take (record, word) slots, a record's
// slots covering the aligned words of its row; a word a row shares with a
// neighbouring row is written by bytes, every other word whole.
__device__ __forceinline__ void put_rows(uint8_t *row0, uint8_t *crow0, int64_t stride, int nrec, const uint8_t *buf,
                                         int64_t n, const int64_t *src, const int32_t *len, uint32_t pad, int pmode,
                                         int cmode, int lane) {
    const int W = (int)((stride + 6) >> 2) + 1;  // aligned words a row can touch, at most
    int j = lane / W, k = lane - j * W;
#pragma unroll 8
    for (int slot = lane; slot < nrec * W; slot += 64) {
        const uintptr_t a0 = (uintptr_t)(row0 + (int64_t)j * stride);
        const uintptr_t wa = (a0 & ~(uintptr_t)3) + 4 * (uintptr_t)k;
        const int64_t x0 = (int64_t)(wa - a0);  // row offset of the word's first byte
        if (x0 < stride) {
            const uint32_t v = row4(buf, n, src[j], len[j], x0, pad);
            put_word(wa, a0, x0, stride, cvt4(v, pmode));
            if (crow0) {
                const uintptr_t d = (uintptr_t)(crow0 - row0);
                put_word(wa + d, a0 + d, x0, stride, cvt4(v, cmode));
            }
        }
        k += 64;
        while (k >= W) {
            k -= W;
            j++;
        }
    }
}

// Chunks of 64 records per wave. A lane per record: line spans, lengths and
// the name span (the first ' ' or '/' of the header, four bytes at a time); the
// first read longer than stride. Then the wave writes the chunk's base,
// converted-base and quality rows.
__global__ __launch_bounds__(256) void fq_measure_kernel(const uint8_t *__restrict__ buf, int64_t n,
                                                         const uint64_t *__restrict__ nl, int64_t max_records,
                                                         int64_t stride, int mode, uint64_t *__restrict__ info,
                                                         uint32_t *__restrict__ lens, uint32_t *__restrict__ qual_lens,
                                                         uint32_t *__restrict__ name_len,
                                                         uint64_t *__restrict__ name_src, uint8_t *__restrict__ rows,
                                                         uint8_t *__restrict__ conv_rows,
                                                         uint8_t *__restrict__ qual_rows) {
    __shared__ FqLds S;
    const int64_t recs = (int64_t)info[0];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t wave = (int64_t)__builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    const int64_t n_waves = (int64_t)gridDim.x * 4;
    // records past the batch: name length 0 (the offset scan reads them)
    for (int64_t r = recs + wave * 64 + lane; r < max_records; r += n_waves * 64) name_len[r] = 0;
    for (int64_t c0 = wave * FQ_CHUNK; c0 < recs; c0 += n_waves * FQ_CHUNK) {
        const int nrec = (int)(recs - c0 < FQ_CHUNK ? recs - c0 : FQ_CHUNK);
        const int64_t r = c0 + lane;
        int64_t ss = 0, se = 0, qs = 0, qe = 0;
        if (lane < nrec) {
            // one batch of loads after the newline positions: the bytes before the line
            // ends ('\r') and the header's first 36 bytes
            const int64_t L0 = 4 * r;
            const int64_t hs = L0 == 0 ? 0 : (int64_t)nl[L0 - 1] + 1;
            const int64_t hn = (int64_t)nl[L0], sn = (int64_t)nl[L0 + 1], pn = (int64_t)nl[L0 + 2];
            const int64_t qn = (int64_t)nl[L0 + 3];
            ss = hn + 1;
            qs = pn + 1;
            const bool hcr = hn > hs && buf[hn - 1] == '\r';
            const bool scr = sn > ss && buf[sn - 1] == '\r';
            const bool qcr = qn > qs && buf[qn - 1] == '\r';
            uint32_t hw[9];
            {
                // ten aligned words from the one holding hs (each holds a byte below n, or
                // is not read), shifted into the nine 4-byte windows from hs
                const uintptr_t a = (uintptr_t)(buf + hs);
                const uint32_t *w = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
                const int64_t lead = (int64_t)(a & 3);  // bytes of word 0 before hs
                uint32_t aw[10];
#pragma unroll
                for (int i = 0; i < 10; i++) aw[i] = hs - lead + 4 * i < n ? w[i] : 0u;
#pragma unroll
                for (int i = 0; i < 9; i++) hw[i] = __builtin_amdgcn_alignbyte(aw[i + 1], aw[i], (uint32_t)lead);
                if (hs + 36 > n) {  // bytes at or past n read as 0 (their words may hold stale bytes)
#pragma unroll
                    for (int i = 0; i < 9; i++) {
                        const int64_t v = n - (hs + 4 * i);
                        hw[i] = v <= 0 ? 0u : v >= 4 ? hw[i] : hw[i] & ((1u << (8 * v)) - 1);
                    }
                }
            }
            const int64_t he = hn - (hcr ? 1 : 0);
            se = sn - (scr ? 1 : 0);
            qe = qn - (qcr ? 1 : 0);
            const int at = (he > hs && (hw[0] & 255u) == '@') ? 1 : 0;  // the name skips a leading '@'
            const int64_t s0 = hs + at;
            int64_t e0 = he;
            bool found = false;
#pragma unroll
            for (int i = 7; i >= 0; i--) {  // the first ' ' or '/' among the name's first 32 bytes
                const uint32_t w = at ? __builtin_amdgcn_alignbyte(hw[i + 1], hw[i], 1u) : hw[i];
                const int64_t q = s0 + 4 * i;
                uint32_t hit = zero_bytes(w ^ 0x20202020u) | zero_bytes(w ^ 0x2F2F2F2Fu);
                if (he - q < 4) hit = he - q > 0 ? hit & ((1u << (8 * (he - q))) - 1) : 0u;
                if (hit) {
                    e0 = q + (__builtin_ctz(hit) >> 3);
                    found = true;
                }
            }
            if (!found) {
                for (int64_t q = s0 + 32; q < he; q += 4) {  // longer names
                    uint32_t w = 0;
                    if (q + 4 <= n) {
                        w = text4(buf, q);
                    } else {
                        for (int j = 0; j < 4 && q + j < n; j++) w |= (uint32_t)buf[q + j] << (8 * j);
                    }
                    uint32_t hit = zero_bytes(w ^ 0x20202020u) | zero_bytes(w ^ 0x2F2F2F2Fu);
                    if (he - q < 4) hit &= (1u << (8 * (he - q))) - 1;
                    if (hit) {
                        e0 = q + (__builtin_ctz(hit) >> 3);
                        break;
                    }
                }
            }
            const int64_t slen = se - ss;
            lens[r] = (uint32_t)slen;
            qual_lens[r] = (uint32_t)(qe - qs);
            name_len[r] = (uint32_t)(e0 - s0);
            name_src[r] = (uint64_t)s0;
            if (slen > stride) atomicMin((unsigned long long *)&info[7], (unsigned long long)r);
        }
        S.src[wv][lane] = ss;
        S.len[wv][lane] = (int32_t)(se - ss);
        wsync();
        put_rows(rows + c0 * stride, conv_rows && (((uintptr_t)conv_rows - (uintptr_t)rows) & 3) == 0
                                         ? conv_rows + c0 * stride : nullptr,
                 stride, nrec, buf, n, S.src[wv], S.len[wv], 0x4E4E4E4Eu, 0, mode, lane);
        if (conv_rows && (((uintptr_t)conv_rows - (uintptr_t)rows) & 3) != 0)
            put_rows(conv_rows + c0 * stride, nullptr, stride, nrec, buf, n, S.src[wv], S.len[wv], 0x4E4E4E4Eu, mode,
                     0, lane);
        wsync();
        S.src[wv][lane] = qs;
        S.len[wv][lane] = (int32_t)(qe - qs);
        wsync();
        if (qual_rows)
            put_rows(qual_rows + c0 * stride, nullptr, stride, nrec, buf, n, S.src[wv], S.len[wv], 0u, 0, 0, lane);
        wsync();
    }
}

// Chunks of 64 records per wave: the chunk's names are one run of the output
// (the offsets are a scan), taken a 4-byte output word per lane (whole-word
// stores inside the run); a word's records come from a binary search over the
// chunk's offsets.
__global__ __launch_bounds__(256) void fq_names_kernel(const uint8_t *__restrict__ buf,
                                                       const uint64_t *__restrict__ info,
                                                       const uint64_t *__restrict__ name_src,
                                                       const int64_t *__restrict__ name_off, int64_t names_cap,
                                                       uint8_t *__restrict__ names) {
    __shared__ FqLds S;
    const int64_t recs = (int64_t)info[0];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t wave = (int64_t)__builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    for (int64_t c0 = wave * FQ_CHUNK; c0 < recs; c0 += (int64_t)gridDim.x * 4 * FQ_CHUNK) {
        const int nrec = (int)(recs - c0 < FQ_CHUNK ? recs - c0 : FQ_CHUNK);
        if (lane < nrec) {
            S.src[wv][lane] = (int64_t)name_src[c0 + lane];
            S.off[wv][lane] = name_off[c0 + lane];
        }
        if (lane == 0) S.off[wv][nrec] = name_off[c0 + nrec];
        wsync();
        const int64_t o0 = S.off[wv][0], o1 = S.off[wv][nrec];
        const int64_t end = o1 < names_cap ? o1 : names_cap;  // names past the capacity are not written
        // a lane per aligned output word of [o0, end): its bytes' records from one binary
        // search (the first byte's) and the offsets after it
        const int64_t w0 = o0 >> 2;
        for (int64_t wd = w0 + lane; 4 * wd < end; wd += 64) {
            const int64_t x0 = 4 * wd;
            const int64_t xs = x0 > o0 ? x0 : o0;  // the word's first byte inside the run
            int lo = 0, hi = nrec - 1;  // the record whose name holds byte xs
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (S.off[wv][mid] <= xs)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            uint32_t v = 0, keep = 0;
            for (int q = 0; q < 4; q++) {
                const int64_t x = x0 + q;
                if (x < xs || x >= end) continue;
                while (lo + 1 < nrec && S.off[wv][lo + 1] <= x) lo++;
                if (S.off[wv][lo + 1] > names_cap) continue;  // this name does not fit whole
                v |= (uint32_t)buf[S.src[wv][lo] + (x - S.off[wv][lo])] << (8 * q);
                keep |= 0xFFu << (8 * q);
            }
            if (keep == 0xFFFFFFFFu && (((uintptr_t)(names + x0)) & 3) == 0) {
                *reinterpret_cast<uint32_t *>(names + x0) = v;
            } else {
                for (int q = 0; q < 4; q++)
                    if ((keep >> (8 * q)) & 1) names[x0 + q] = (uint8_t)(v >> (8 * q));
            }
        }
        wsync();
    }
}

// The batch stops at its first read longer than stride (status 4) or whose
// name passes names_cap (status 5), as a sequential reader would: the records
// before it stand, and the bytes consumed end with them.
__global__ void fq_finish_kernel(uint64_t *__restrict__ info, const uint64_t *__restrict__ nl,
                                 const int64_t *__restrict__ name_off, int64_t names_cap, int64_t n) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint64_t recs = info[0];
    uint64_t stop = recs;
    uint32_t why = 0;
    if (info[7] < stop) {
        stop = info[7];
        why = FQ_TOO_LONG;
    }
    if (recs && (uint64_t)name_off[recs] > (uint64_t)names_cap) {
        uint64_t lo = 0, hi = recs - 1;  // first r with name_off[r + 1] > names_cap
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            if ((uint64_t)name_off[mid + 1] > (uint64_t)names_cap)
                hi = mid;
            else
                lo = mid + 1;
        }
        if (lo < stop) {
            stop = lo;
            why = FQ_NAMES;
        }
    }
    if (stop < recs) {
        info[0] = stop;
        info[3] = why;
        const uint64_t consumed = stop ? nl[4 * stop - 1] + 1 : 0;
        info[1] = consumed > (uint64_t)n ? (uint64_t)n : consumed;
    }
    info[4] = info[0] ? (uint64_t)name_off[info[0]] : 0;
}

// Longest read: a block-wide maximum, then one atomic per block (a small grid).
__global__ __launch_bounds__(256) void fq_maxlen_kernel(const uint32_t *__restrict__ lens, uint64_t *__restrict__ info) {
    __shared__ uint32_t sh[4];
    const int64_t recs = (int64_t)info[0];
    uint32_t m = 0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < recs; r += (int64_t)gridDim.x * blockDim.x)
        m = max(m, lens[r]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = max(max(sh[0], sh[1]), max(sh[2], sh[3]));
        if (m) atomicMax((unsigned long long *)&info[2], (unsigned long long)m);
    }
}

struct FqScratch {
    size_t counts, tile_base, partial, total_nl, nl, name_len, name_src, name_part, total;
};

FqScratch fq_layout(int64_t n, int64_t max_records) {
    const int64_t tiles = (n + FQ_TILE - 1) / FQ_TILE, t1 = tiles > 0 ? tiles : 1;
    const int64_t r1 = max_records > 0 ? max_records : 1;
    FqScratch s;
    size_t o = 0;
    auto take = [&](size_t b) {
        const size_t r = o;
        o += (b + 255) & ~(size_t)255;
        return r;
    };
    s.counts = take((size_t)t1 * 4);
    s.tile_base = take((size_t)(t1 + 1) * 8);
    s.partial = take(scan_scratch_bytes(t1 > r1 ? t1 : r1));
    s.total_nl = take(8);
    s.nl = take((size_t)(4 * r1 + 4) * 8);
    s.name_len = take((size_t)r1 * 4);
    s.name_src = take((size_t)r1 * 8);
    s.name_part = take(scan_scratch_bytes(r1));
    s.total = o;
    return s;
}

}  // namespace

extern "C" {

int64_t ma_fastq_scratch_bytes(int64_t nbytes, int64_t max_records) {
    return (int64_t)fq_layout(nbytes < 0 ? 0 : nbytes, max_records).total;
}

int ma_fastq_parse_dev(void *stream, const uint8_t *d_buf, int64_t nbytes, int32_t final, int64_t max_records,
                       int32_t stride, int32_t bs_mode, uint8_t *d_rows, uint8_t *d_conv_rows, uint8_t *d_qual_rows,
                       uint32_t *d_lens, uint32_t *d_qual_lens, uint8_t *d_names, int64_t names_cap,
                       int64_t *d_name_off, uint64_t *d_info, void *d_scratch, int64_t scratch_bytes) {
    MA_REQUIRE(nbytes >= 0 && (nbytes == 0 || d_buf) && max_records >= 1 && stride >= 1,
               "ma_fastq_parse_dev: bad arguments");
    MA_REQUIRE(d_rows && d_lens && d_qual_lens && d_names && d_name_off && d_info, "ma_fastq_parse_dev: null output");
    MA_REQUIRE(bs_mode >= 0 && bs_mode <= 2, "ma_fastq_parse_dev: conversion must be none, C2T (1) or G2A (2)");
    MA_REQUIRE(max_records <= 0x7FFFFFFFll, "ma_fastq_parse_dev: max_records too large");
    const FqScratch L = fq_layout(nbytes, max_records);
    MA_REQUIRE(scratch_bytes >= (int64_t)L.total, "ma_fastq_parse_dev: scratch too small");
    hipStream_t st = (hipStream_t)stream;
    char *s = (char *)d_scratch;
    const int64_t tiles = (nbytes + FQ_TILE - 1) / FQ_TILE;
    uint32_t *counts = (uint32_t *)(s + L.counts);
    uint64_t *tile_base = (uint64_t *)(s + L.tile_base), *total_nl = (uint64_t *)(s + L.total_nl);
    uint64_t *nl = (uint64_t *)(s + L.nl);
    const int64_t cap = 4 * max_records + 4;
    if (tiles > 0) {
        const unsigned grid = (unsigned)(tiles < 8192 ? tiles : 8192);
        fq_count_kernel<<<grid, FQ_TB, 0, st>>>(d_buf, nbytes, tiles, counts);
        MA_LAUNCHED();
        MA_TRY(exclusive_scan_u32(st, tiles, counts, tile_base, (uint64_t *)(s + L.partial), total_nl));
        fq_lines_kernel<<<grid, FQ_TB, 0, st>>>(d_buf, nbytes, tiles, tile_base, nl, cap);
        MA_LAUNCHED();
    } else {
        MA_TRY(fill_u64(st, total_nl, 1, 0));
    }
    fq_settle_kernel<<<1, 64, 0, st>>>(d_buf, nbytes, final, max_records, nl, cap, total_nl, d_info);
    MA_LAUNCHED();
    uint32_t *name_len = (uint32_t *)(s + L.name_len);
    uint64_t *name_src = (uint64_t *)(s + L.name_src);
    const unsigned g = (unsigned)grid_for(max_records);
    const int64_t chunks = (max_records + FQ_CHUNK - 1) / FQ_CHUNK;  // a wave per chunk of records
    const unsigned gm = (unsigned)((chunks + 3) / 4 < 65536 ? (chunks + 3) / 4 : 65536);
    fq_measure_kernel<<<gm, 256, 0, st>>>(d_buf, nbytes, nl, max_records, stride, bs_mode, d_info, d_lens,
                                          d_qual_lens, name_len, name_src, d_rows, d_conv_rows, d_qual_rows);
    MA_LAUNCHED();
    MA_TRY(exclusive_scan_u32(st, max_records, name_len, (uint64_t *)d_name_off, (uint64_t *)(s + L.name_part),
                              (uint64_t *)d_name_off + max_records));
    fq_finish_kernel<<<1, 64, 0, st>>>(d_info, nl, d_name_off, names_cap, nbytes);
    MA_LAUNCHED();
    fq_maxlen_kernel<<<g < 1024u ? g : 1024u, 256, 0, st>>>(d_lens, d_info);
    MA_LAUNCHED();
    const int64_t blocks = gm;
    fq_names_kernel<<<(unsigned)blocks, 256, 0, st>>>(d_buf, d_info, name_src, d_name_off, names_cap, d_names);
    MA_LAUNCHED();
    return MA_OK;
}

}  // extern "C"
