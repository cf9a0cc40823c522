// BGZF members compressed on the device.
//
// The stream is cut into consecutive 65280-byte blocks exactly as the
// reference's writer cuts it (fq2bam-meth/python/bam_emit.py:103-185); each
// block becomes one gzip member with the BC extra field, a raw deflate body,
// CRC32 and ISIZE (_bgzf_block, :77-100). The deflate body is this build's own
// compressor, so the compressed bytes differ from zlib's while the members
// decode to the same stream:
//
//   * one 1024-thread workgroup per block (persistent grid), the block staged
//     in LDS (64 bytes per thread);
//   * LZ77: a 4-byte hash table in LDS filled chunk by chunk (1024 positions
//     per step, in order), so each position gets the latest earlier-chunk
//     position with its hash as a match candidate, plus distance 1 for runs;
//     each thread parses its 64-byte segment greedily (matches of 4..64 bytes
//     stay inside the segment, so the parses are independent);
//   * one dynamic-Huffman block (RFC 1951 3.2.7) per member: literal/length
//     and distance code lengths from two-queue Huffman builds (two threads at
//     once) limited to 15 bits (7 for the code-length code) by Kraft-sum
//     repair, canonical codes; a stored block when that is not smaller;
//   * per-thread bit offsets from a block scan; each thread writes its
//     segment's codes straight into the block's output slot (plain stores for
//     the words it owns, atomic OR for the two it may share, which alone are
//     cleared first); tokens live in a scratch laid out so a wave's accesses
//     are 64 consecutive words;
//   * CRC32: per-thread CRCs of the 64-byte segments joined pairwise by the
//     GF(2) shift crc(A||B) = crc(A) * x^(8|B|) + crc(B) mod P;
//   * a second kernel writes every member (18-byte header, body, trailer) at
//     its offset from a scan of the member sizes, and the EOF member.
#include "ma_sort.hpp"

using namespace ma;

namespace {

constexpr int DF_T = 1024;            // threads per workgroup
constexpr int DF_SEG = 64;            // input bytes per thread
constexpr int DF_IN = 65280;          // bytes per BGZF block
constexpr int DF_OUT_WORDS = 16384;   // 64 KiB output slot per block
constexpr int64_t DF_SLOT = 65536;    // staging bytes per block
constexpr int NLIT = 286;             // literal/length alphabet
constexpr int NDC = 30;               // distance alphabet
constexpr int NCL = 19;               // code-length alphabet
constexpr int NSEQ = NLIT + NDC;      // code-length sequence entries at most
constexpr int HASH_BITS = 13;         // match-candidate table: 8192 entries
constexpr uint32_t CRC_POLY = 0xEDB88320u;  // reflected CRC-32 polynomial
constexpr uint32_t STORED = 0xFFFFFFFFu;

// CL alphabet transmission order (RFC 1951 3.2.7)
__device__ __forceinline__ int cl_order(int i) {
    constexpr uint32_t lo = 16 | 17 << 5 | 18 << 10 | 0 << 15 | 8 << 20 | 7 << 25;  // entries 0..5
    constexpr uint32_t mid = 9 | 6 << 5 | 10 << 10 | 5 << 15 | 11 << 20 | 4 << 25;  // 6..11
    constexpr uint32_t hi = 12 | 3 << 5 | 13 << 10 | 2 << 15 | 14 << 20 | 1 << 25;  // 12..17
    if (i < 6) return (lo >> (5 * i)) & 31;
    if (i < 12) return (mid >> (5 * (i - 6))) & 31;
    if (i < 18) return (hi >> (5 * (i - 12))) & 31;
    return 15;
}

// a * b mod P in the reflected representation (bit 31 = x^0).
__device__ __forceinline__ uint32_t multmodp(uint32_t a, uint32_t b) {
    uint32_t p = 0;
#pragma unroll 4
    for (int i = 0; i < 32; i++) {
        if (a & (0x80000000u >> i)) p ^= b;
        b = (b & 1) ? (b >> 1) ^ CRC_POLY : b >> 1;
    }
    return p;
}

// x^(8 len) mod P from the x^(2^k) table.
__device__ __forceinline__ uint32_t x8n(const uint32_t *xpow, uint32_t len) {
    uint32_t r = 0x80000000u;  // x^0
    for (int k = 0; len; k++, len >>= 1)
        if (len & 1) r = multmodp(xpow[k + 3], r);
    return r;
}

__device__ __forceinline__ uint32_t bitrev(uint32_t c, int len) { return __brev(c) >> (32 - len); }

// Code lengths (<= maxlen) for the m used symbols sorted[0..m) (ascending
// (freq, symbol)); lengths of unused symbols stay 0. m >= 1. A single used
// symbol gets length 1 and a second symbol `spare` length 1 too, so every
// code is complete (inflate rejects incomplete code-length codes).
__device__ void huff_lengths(const uint32_t *freq, const uint16_t *sorted, int m, int maxlen, int spare,
                             uint8_t *len, uint32_t *nf, uint16_t *par, uint8_t *dep, uint32_t *blc) {
    if (m == 1) {
        len[sorted[0]] = 1;
        len[sorted[0] == spare ? spare + 1 : spare] = 1;
        return;
    }
    // two-queue Huffman: leaves 0..m-1 in sorted order, internal node k has id m + k
    int i = 0, j = 0;
    for (int k = 0; k < m - 1; k++) {
        uint32_t f = 0;
        for (int pick = 0; pick < 2; pick++) {
            const bool leaf = i < m && (j >= k || freq[sorted[i]] <= nf[j]);
            if (leaf) {
                par[i] = (uint16_t)(m + k);
                f += freq[sorted[i]];
                i++;
            } else {
                par[m + j] = (uint16_t)(m + k);
                f += nf[j];
                j++;
            }
        }
        nf[k] = f;
    }
    dep[2 * m - 2] = 0;
    for (int node = 2 * m - 3; node >= 0; node--) {
        const int d = dep[par[node]] + 1;
        dep[node] = (uint8_t)(d > 255 ? 255 : d);
    }
    for (int l = 0; l < 32; l++) blc[l] = 0;
    for (int s = 0; s < m; s++) blc[dep[s] > maxlen ? maxlen : dep[s]]++;
    // Kraft repair after clamping: sum 2^(maxlen - l) over leaves must equal 2^maxlen
    uint32_t kraft = 0;
    for (int l = 1; l <= maxlen; l++) kraft += blc[l] << (maxlen - l);
    const uint32_t target = 1u << maxlen;
    // (the inner searches stop at length 1: every step below is bounded)
    while (kraft > target) {  // lengthen the deepest code that can still grow
        int l = maxlen - 1;
        while (l > 1 && blc[l] == 0) l--;
        if (blc[l] == 0) break;
        blc[l]--;
        blc[l + 1]++;
        kraft -= 1u << (maxlen - l - 1);
    }
    while (kraft < target) {  // shorten the deepest code: its 2^(maxlen-l) divides the deficit
        int l = maxlen;
        while (l > 2 && blc[l] == 0) l--;
        if (blc[l] == 0) break;
        blc[l]--;
        blc[l - 1]++;
        kraft += 1u << (maxlen - l);
    }
    int s = 0;  // least frequent symbols take the longest codes
    for (int l = maxlen; l >= 1; l--)
        for (uint32_t c = 0; c < blc[l]; c++) len[sorted[s++]] = (uint8_t)l;
}

// Canonical codes (bit-reversed for LSB-first output) from lengths; blc has
// 32 entries: counts in 0..15, next codes in 16..31 (shared memory).
__device__ void canonical(const uint8_t *len, int nsym, int maxlen, uint16_t *code, uint32_t *blc) {
    for (int l = 0; l < 32; l++) blc[l] = 0;
    for (int s = 0; s < nsym; s++) blc[len[s]]++;
    blc[0] = 0;
    uint32_t c = 0;
    for (int l = 1; l <= maxlen; l++) {
        c = (c + blc[l - 1]) << 1;
        blc[16 + l] = c;
    }
    for (int s = 0; s < nsym; s++)
        if (len[s]) code[s] = (uint16_t)bitrev(blc[16 + len[s]]++, len[s]);
}

// LSB-first bit writer into the shared buffer: words strictly inside the
// writer's range are stored, the first and last (possibly shared) are ORed.
struct BitWriter {
    uint32_t *out;
    uint32_t wi;
    uint64_t acc;
    int nb;
    bool first;
    __device__ BitWriter(uint32_t *o, uint32_t pos) : out(o), wi(pos >> 5), acc(0), nb((int)(pos & 31)), first(true) {}
    __device__ __forceinline__ void put(uint32_t v, int n) {
        acc |= (uint64_t)v << nb;
        nb += n;
        if (nb >= 32) {
            if (first)
                atomicOr(&out[wi], (uint32_t)acc);
            else
                out[wi] = (uint32_t)acc;
            first = false;
            wi++;
            acc >>= 32;
            nb -= 32;
        }
    }
    __device__ __forceinline__ void finish() {
        if (nb > 0) atomicOr(&out[wi], (uint32_t)acc);
    }
};

__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t *wsum, uint32_t &total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int k = 0; k < DF_T / 64; k++) {
        const uint32_t s = wsum[k];
        if (k < w) pre += s;
        tot += s;
    }
    total = tot;
    __syncthreads();
    return pre + x - v;
}

struct DfShared {
    uint8_t data[DF_IN];
    uint32_t head[1 << HASH_BITS];  // per 4-byte hash: 1 + the latest position of earlier chunks
    uint32_t crc_tab[256];
    uint32_t xpow[40];         // x^(2^k) mod P (reflected), k = 0..39
    uint32_t shift64[DF_T];    // x^(8 * 64 k) mod P: the CRC shift over k whole segments
    uint32_t crcw[DF_T / 64];  // per-wave XOR of the shifted segment CRCs
    uint32_t sf[NLIT];         // frequencies in ranked order
    uint32_t dsf[NDC];
    uint32_t clblc[32];
    uint32_t part[DF_T];       // per-thread CRCs, then bit counts
    uint32_t wsum[DF_T / 64];
    uint8_t ntok[DF_T];        // tokens of each thread's segment
    uint32_t freq[NLIT];
    uint32_t dfreq[NDC];
    uint8_t lit_len[NLIT];
    uint16_t lit_code[NLIT];   // bit-reversed canonical codes
    uint8_t d_len[NDC];
    uint16_t d_code[NDC];
    uint16_t sorted[NLIT];     // used symbols by ascending (freq, symbol)
    uint16_t dsorted[NDC];
    uint32_t nf[NLIT];         // Huffman build scratch (literal/length and code-length trees)
    uint16_t par[2 * NLIT];
    uint8_t dep[2 * NLIT];
    uint32_t dnf[NDC];         // Huffman build scratch (distance tree, built concurrently)
    uint16_t dpar[2 * NDC];
    uint8_t ddep[2 * NDC];
    uint32_t dblc[32];
    uint8_t cl_sym[NSEQ];
    uint8_t cl_extra[NSEQ];
    uint32_t cl_freq[NCL];
    uint8_t cl_len[NCL];
    uint16_t cl_code[NCL];
    uint16_t cl_sorted[NCL];
    uint32_t blc[32];
    uint32_t misc[8];          // 0 used lit/len symbols, 1 CL entries, 2 last used lit/len symbol + 1,
                               // 3 hclen, 5 used distance symbols, 6 last used distance symbol + 1,
                               // 7 used CL symbols
};

// Length symbol (257..285) and extra bits of a match length 3..258 (RFC 1951 3.2.5).
__device__ __forceinline__ void len_code(uint32_t L, uint32_t &sym, uint32_t &ebits, uint32_t &extra) {
    if (L == 258) {
        sym = 285;
        ebits = extra = 0;
        return;
    }
    const uint32_t v = L - 3;
    if (v < 8) {
        sym = 257 + v;
        ebits = extra = 0;
        return;
    }
    const uint32_t e = 31 - __clz(v) - 2;
    sym = 257 + 4 * (e + 1) + ((v >> e) & 3);
    ebits = e;
    extra = v & ((1u << e) - 1);
}

// Distance symbol (0..29) and extra bits of a distance 1..32768.
__device__ __forceinline__ void dist_code(uint32_t D, uint32_t &sym, uint32_t &ebits, uint32_t &extra) {
    const uint32_t d = D - 1;
    if (d < 4) {
        sym = d;
        ebits = extra = 0;
        return;
    }
    const uint32_t e = 31 - __clz(d) - 1;
    sym = 2 * (e + 1) + ((d >> e) & 1);
    ebits = e;
    extra = d & ((1u << e) - 1);
}

__device__ __forceinline__ uint32_t hash4(const uint8_t *d, int p) {
    const uint32_t x = (uint32_t)d[p] | (uint32_t)d[p + 1] << 8 | (uint32_t)d[p + 2] << 16 | (uint32_t)d[p + 3] << 24;
    return (x * 2654435761u) >> (32 - HASH_BITS);
}

__device__ __forceinline__ int match_len(const uint8_t *d, int q, int p, int maxl) {
    int i = 0;
    while (i < maxl && d[q + i] == d[p + i]) i++;
    return i;
}

// Ranks the used symbols of freq[0..nsym) by (freq, symbol) into sorted[];
// threads [t0, t0 + nsym) take one symbol each; counts them into *m.
__device__ __forceinline__ void rank_symbols(const uint32_t *freq, int nsym, uint16_t *sorted, uint32_t *sf,
                                             uint32_t *m, int t0) {
    const int s = (int)threadIdx.x - t0;
    if (s >= 0 && s < nsym) {
        const uint32_t f = freq[s];
        if (f) {
            int r = 0;
            for (int u = 0; u < nsym; u++) {
                const uint32_t g = freq[u];
                r += g && (g < f || (g == f && u < s));
            }
            sorted[r] = (uint16_t)s;
            sf[r] = f;
            atomicAdd(m, 1u);
        }
    }
}

// Two-queue Huffman over m >= 2 leaves with ascending weights sf[0..m):
// internal node k (id m + k) gets weight nf[k]; par[] = parent id of every
// node (leaves first). One thread.
__device__ void huff_tree(const uint32_t *sf, int m, uint32_t *nf, uint16_t *par) {
    int i = 0, j = 0;
    uint32_t li = sf[0];
    for (int k = 0; k < m - 1; k++) {
        uint32_t f = 0;
        for (int pick = 0; pick < 2; pick++) {
            if (i < m && (j >= k || li <= nf[j])) {
                par[i] = (uint16_t)(m + k);
                f += li;
                i++;
                li = i < m ? sf[i] : 0;
            } else {
                par[m + j] = (uint16_t)(m + k);
                f += nf[j];
                j++;
            }
        }
        nf[k] = f;
    }
}

// Depth of leaf x: steps up to the root (id root).
__device__ __forceinline__ int leaf_depth(const uint16_t *par, int x, int root) {
    int d = 0;
    while (x != root && d < 2 * NLIT) {
        x = par[x];
        d++;
    }
    return d;
}

// Per-length counts blc[1..maxlen] (deeper leaves already counted at maxlen)
// repaired so that the Kraft sum is exactly 1. One thread.
__device__ void kraft_repair(uint32_t *blc, int maxlen) {
    uint32_t kraft = 0;
    for (int l = 1; l <= maxlen; l++) kraft += blc[l] << (maxlen - l);
    const uint32_t target = 1u << maxlen;
    while (kraft > target) {  // lengthen the deepest code that can still grow
        int l = maxlen - 1;
        while (l > 1 && blc[l] == 0) l--;
        if (blc[l] == 0) break;
        blc[l]--;
        blc[l + 1]++;
        kraft -= 1u << (maxlen - l - 1);
    }
    while (kraft < target) {  // shorten the deepest code: its 2^(maxlen-l) divides the deficit
        int l = maxlen;
        while (l > 2 && blc[l] == 0) l--;
        if (blc[l] == 0) break;
        blc[l]--;
        blc[l - 1]++;
        kraft += 1u << (maxlen - l);
    }
}

// Length of the r-th least frequent symbol: lengths handed out longest first.
__device__ __forceinline__ uint8_t rank_length(const uint32_t *blc, int maxlen, int r) {
    uint32_t acc = 0;
    for (int l = maxlen; l >= 1; l--) {
        acc += blc[l];
        if ((uint32_t)r < acc) return (uint8_t)l;
    }
    return 1;
}

// Canonical code of symbol s (bit-reversed): the first code of its length
// plus the number of smaller symbols of that length.
__device__ __forceinline__ uint16_t canon_code(const uint8_t *len, const uint32_t *blc, int s) {
    const int l = len[s];
    if (l == 0) return 0;
    uint32_t code = 0;
    for (int k = 1; k <= l; k++) code = (code + (k > 1 ? blc[k - 1] : 0)) << 1;
    for (int u = 0; u < s; u++) code += len[u] == l;
    return (uint16_t)bitrev(code, l);
}

// Code-length symbols of a run of `run` equal lengths v (RFC 1951 3.2.7:
// 16 repeats the previous length 3..6 times, 17 / 18 give 3..10 / 11..138
// zeros); writes them when sym is not null; returns their number.
__device__ int rle_emit(int v, int run, uint8_t *sym, uint8_t *extra) {
    int ns = 0;
    auto put = [&](int sv, int ev) {
        if (sym) {
            sym[ns] = (uint8_t)sv;
            extra[ns] = (uint8_t)ev;
        }
        ns++;
    };
    if (v == 0) {
        while (run >= 11) {
            const int r = run < 138 ? run : 138;
            put(18, r - 11);
            run -= r;
        }
        if (run >= 3) {
            put(17, run - 3);
            run = 0;
        }
    } else {
        put(v, 0);
        run--;
        while (run >= 3) {
            const int r = run < 6 ? run : 6;
            put(16, r - 3);
            run -= r;
        }
    }
    for (; run > 0; run--) put(v, 0);
    return ns;
}

// Tokens: a literal byte b is b; a match is 1 << 31 | length << 16 | distance.
constexpr uint32_t TOK_MATCH = 1u << 31;
constexpr int MIN_MATCH = 4;

__global__ __launch_bounds__(DF_T) void bgzf_deflate_kernel(const uint8_t *__restrict__ src, int64_t n,
                                                            int64_t nb, int level,
                                                            uint32_t *__restrict__ slots,
                                                            uint32_t *__restrict__ body_len,
                                                            uint32_t *__restrict__ crc_out,
                                                            uint32_t *__restrict__ member_bytes,
                                                            uint16_t *__restrict__ cand_all,
                                                            uint32_t *__restrict__ tok_all) {
    __shared__ DfShared S;
    const int t = threadIdx.x;
    const bool huffman = level > 0, lz = level > 0;
    uint16_t *const cand = cand_all + (int64_t)blockIdx.x * DF_IN;        // this workgroup's scratch
    // token k of thread t at tok[k * DF_T]: a wave's k-th tokens are 64 consecutive words
    uint32_t *const tok = tok_all + (int64_t)blockIdx.x * DF_T * DF_SEG + t;
    if (t < 256) {
        uint32_t c = (uint32_t)t;
        for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ CRC_POLY : c >> 1;
        S.crc_tab[t] = c;
    }
    if (t == 0) {
        uint32_t x = 0x40000000u;  // x^1
        for (int k = 0; k < 40; k++) {
            S.xpow[k] = x;
            x = multmodp(x, x);
        }
    }
    __syncthreads();
    S.shift64[t] = x8n(S.xpow, (uint32_t)(DF_SEG * t));
    __syncthreads();
    for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t base = b * DF_IN;
        const int bl = (int)(n - base < DF_IN ? n - base : DF_IN);
        uint32_t *const slot = slots + b * (DF_SLOT / 4);
        // 1. stage the block, clear the tables and the block's output slot
        for (int i = t; i < (1 << HASH_BITS); i += DF_T) S.head[i] = 0;
        for (int i = t; i < NLIT; i += DF_T) {
            S.freq[i] = 0;
            S.lit_len[i] = 0;
        }
        if (t < NDC) {
            S.dfreq[t] = 0;
            S.d_len[t] = 0;
        }
        if (t < NCL) {
            S.cl_freq[t] = 0;
            S.cl_len[t] = 0;
        }
        if (t < 8) S.misc[t] = 0;
        const uint8_t *blk = src + base;
        for (int i = 4 * t; i < bl; i += 4 * DF_T) {
            if (i + 4 <= bl) {
                *reinterpret_cast<uint32_t *>(&S.data[i]) = *reinterpret_cast<const uint32_t *>(blk + i);
            } else {
                for (int k = i; k < bl; k++) S.data[k] = blk[k];
            }
        }
        __syncthreads();
        // 2. per-segment CRC
        const int s0 = t * DF_SEG, s1 = min(bl, s0 + DF_SEG);
        uint32_t c = 0xFFFFFFFFu;
        for (int i = s0; i < s1; i++) c = S.crc_tab[(c ^ S.data[i]) & 255] ^ (c >> 8);
        {
            // crc(S_0 .. S_k) = XOR over segments of crc(S_t) * x^(8 * bytes after S_t) mod P
            uint32_t v = 0;
            if (s0 < bl) {
                const int after = bl - s1;
                v = multmodp(after % DF_SEG == 0 ? S.shift64[after / DF_SEG] : x8n(S.xpow, (uint32_t)after), ~c);
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v ^= __shfl_xor(v, o, 64);
            if ((t & 63) == 0) S.crcw[t >> 6] = v;
        }
        if (t == 0) S.freq[256] = 1;  // end of block
        // 3. match candidates: chunks of DF_T consecutive positions in order;
        //    a position sees the latest earlier-chunk position with its hash
        if (lz) {
            for (int c0 = 0; c0 < bl; c0 += DF_T) {
                const int p = c0 + t;
                const bool has = p + MIN_MATCH <= bl;
                const uint32_t h = has ? hash4(S.data, p) : 0;
                if (p < bl) cand[p] = has ? (uint16_t)S.head[h] : (uint16_t)0;
                __syncthreads();
                if (has) atomicMax(&S.head[h], (uint32_t)(p + 1));
                __syncthreads();
            }
        } else {
            __syncthreads();
        }
        // 4. CRC of the block (the barriers above ordered the per-wave parts)
        uint32_t crc = 0;
        for (int w = 0; w < DF_T / 64; w++) crc ^= S.crcw[w];
        uint32_t nbytes = 0xFFFFFFFFu;
        if (huffman) {
            // 5. greedy parse of the segment (matches stay inside it), symbol counts
            int k = 0;
            for (int p = s0; p < s1;) {
                int best = 0, bd = 0;
                const int maxl = min(258, s1 - p);
                if (lz && maxl >= MIN_MATCH) {
                    const int cq = cand[p];
                    if (cq && p - (cq - 1) <= 32768) {
                        const int l = match_len(S.data, cq - 1, p, maxl);
                        if (l > best) {
                            best = l;
                            bd = p - (cq - 1);
                        }
                    }
                    if (p >= 1 && best < maxl) {  // runs: distance 1
                        const int l = match_len(S.data, p - 1, p, maxl);
                        if (l > best) {
                            best = l;
                            bd = 1;
                        }
                    }
                }
                uint32_t tk;
                if (best >= MIN_MATCH) {
                    uint32_t ls, le, lx, ds, de, dx;
                    len_code((uint32_t)best, ls, le, lx);
                    dist_code((uint32_t)bd, ds, de, dx);
                    atomicAdd(&S.freq[ls], 1u);
                    atomicAdd(&S.dfreq[ds], 1u);
                    tk = TOK_MATCH | (uint32_t)best << 16 | (uint32_t)bd;
                    p += best;
                } else {
                    tk = S.data[p];
                    atomicAdd(&S.freq[tk], 1u);
                    p += 1;
                }
                tok[(k++) * DF_T] = tk;
            }
            S.ntok[t] = (uint8_t)k;
            __syncthreads();
            // 6. the two codes: symbols ranked in parallel, the two-queue merges by
            //    threads 0 and 64 at once, then depths, lengths and codes in parallel
            rank_symbols(S.freq, NLIT, S.sorted, S.sf, &S.misc[0], 0);
            rank_symbols(S.dfreq, NDC, S.dsorted, S.dsf, &S.misc[5], 512);
            if (t < 32) {
                S.blc[t] = 0;
                S.dblc[t] = 0;
                S.clblc[t] = 0;
            }
            __syncthreads();
            const int m = (int)S.misc[0], md = (int)S.misc[5];  // m >= 2: a byte and end of block
            if (t == 0) huff_tree(S.sf, m, S.nf, S.par);
            if (t == 64 && md >= 2) huff_tree(S.dsf, md, S.dnf, S.dpar);
            __syncthreads();
            if (t < m) atomicAdd(&S.blc[min(leaf_depth(S.par, t, 2 * m - 2), 15)], 1u);
            if (t >= 512 && t - 512 < md && md >= 2)
                atomicAdd(&S.dblc[min(leaf_depth(S.dpar, t - 512, 2 * md - 2), 15)], 1u);
            __syncthreads();
            if (t == 0) kraft_repair(S.blc, 15);
            if (t == 64) {
                if (md >= 2) {
                    kraft_repair(S.dblc, 15);
                } else {  // no match: two 1-bit codes; one distance symbol: it and a spare
                    if (md == 0) {
                        S.d_len[0] = S.d_len[1] = 1;
                    } else {
                        const int s = S.dsorted[0];
                        S.d_len[s] = 1;
                        S.d_len[s == 0 ? 1 : 0] = 1;
                    }
                    S.dblc[1] = 2;
                }
            }
            __syncthreads();
            if (t < m) S.lit_len[S.sorted[t]] = rank_length(S.blc, 15, t);
            if (t >= 512 && t - 512 < md && md >= 2) S.d_len[S.dsorted[t - 512]] = rank_length(S.dblc, 15, t - 512);
            __syncthreads();
            if (t < NLIT) {
                S.lit_code[t] = canon_code(S.lit_len, S.blc, t);
                if (S.lit_len[t]) atomicMax(&S.misc[2], (uint32_t)(t + 1));
            }
            if (t >= 512 && t - 512 < NDC) {
                S.d_code[t - 512] = canon_code(S.d_len, S.dblc, t - 512);
                if (S.d_len[t - 512]) atomicMax(&S.misc[6], (uint32_t)(t - 511));
            }
            __syncthreads();
            // code-length sequence (literal/length then distance lengths): maximal runs
            // found in parallel, each coded at its scanned offset
            const int hlit = max(257, (int)S.misc[2]), hdist = max(1, (int)S.misc[6]);
            const int total = hlit + hdist;
            int rv = 0, run = 0;
            uint32_t cnt = 0;
            if (t < total) {
                rv = t < hlit ? S.lit_len[t] : S.d_len[t - hlit];
                const int pv = t == 0 ? -1 : (t - 1 < hlit ? S.lit_len[t - 1] : S.d_len[t - 1 - hlit]);
                if (rv != pv) {
                    run = 1;
                    while (t + run < total && (t + run < hlit ? S.lit_len[t + run] : S.d_len[t + run - hlit]) == rv) run++;
                    cnt = (uint32_t)rle_emit(rv, run, nullptr, nullptr);
                }
            }
            uint32_t ns;
            const uint32_t co = block_excl_scan_1024(cnt, S.wsum, ns);
            if (cnt) rle_emit(rv, run, S.cl_sym + co, S.cl_extra + co);
            __syncthreads();
            for (uint32_t q = (uint32_t)t; q < ns; q += DF_T) atomicAdd(&S.cl_freq[S.cl_sym[q]], 1u);
            __syncthreads();
            rank_symbols(S.cl_freq, NCL, S.cl_sorted, S.sf, &S.misc[7], 0);
            __syncthreads();
            if (t == 0) {
                huff_lengths(S.cl_freq, S.cl_sorted, (int)S.misc[7], 7, 0, S.cl_len, S.nf, S.par, S.dep, S.clblc);
                canonical(S.cl_len, NCL, 7, S.cl_code, S.clblc);
                int hclen = NCL;
                while (hclen > 4 && S.cl_len[cl_order(hclen - 1)] == 0) hclen--;
                S.misc[1] = ns;
                S.misc[3] = (uint32_t)hclen;
            }
            __syncthreads();
            uint32_t hpart = 0;
            for (uint32_t q = (uint32_t)t; q < ns; q += DF_T) {
                const int sym = S.cl_sym[q];
                hpart += S.cl_len[sym] + (sym == 16 ? 2 : sym == 17 ? 3 : sym == 18 ? 7 : 0);
            }
            uint32_t hcl;
            block_excl_scan_1024(hpart, S.wsum, hcl);
            const uint32_t hb_total = 3 + 5 + 5 + 4 + 3 * S.misc[3] + hcl;
            // 7. bit offsets of the segments
            const int ntok = S.ntok[t];
            uint32_t bits = 0;
            for (int q = 0; q < ntok; q++) {
                const uint32_t tk = tok[q * DF_T];
                if (tk & TOK_MATCH) {
                    uint32_t ls, le, lx, ds, de, dx;
                    len_code((tk >> 16) & 0x1FF, ls, le, lx);
                    dist_code(tk & 0xFFFF ? tk & 0xFFFF : 65536, ds, de, dx);
                    bits += S.lit_len[ls] + le + S.d_len[ds] + de;
                } else {
                    bits += S.lit_len[tk];
                }
            }
            uint32_t dbits;
            const uint32_t off = block_excl_scan_1024(bits, S.wsum, dbits);
            const uint32_t hb = hb_total;
            const uint32_t total_bits = hb + dbits + S.lit_len[256];
            nbytes = (total_bits + 7) >> 3;
            if (nbytes < (uint32_t)bl + 5) {
                // 8. header (thread 0), then every segment's codes, EOB after the last one.
                //    Only the first and last word of each writer's bit range can be shared
                //    (ORed by two writers): those are cleared first; the others are stored whole.
                const uint32_t a = hb + off, e = a + bits + (s1 == bl ? S.lit_len[256] : 0u);
                if (s0 < bl) {
                    slot[a >> 5] = 0;
                    slot[(e - 1) >> 5] = 0;
                }
                if (t == 0) {
                    slot[0] = 0;
                    slot[(hb - 1) >> 5] = 0;
                }
                __syncthreads();
                if (t == 0) {
                    BitWriter w(slot, 0);
                    w.put(1, 1);  // BFINAL
                    w.put(2, 2);  // BTYPE = dynamic Huffman
                    w.put((uint32_t)hlit - 257, 5);
                    w.put((uint32_t)hdist - 1, 5);
                    w.put(S.misc[3] - 4, 4);
                    for (uint32_t q = 0; q < S.misc[3]; q++) w.put(S.cl_len[cl_order((int)q)], 3);
                    for (uint32_t q = 0; q < S.misc[1]; q++) {
                        const int sym = S.cl_sym[q];
                        w.put(S.cl_code[sym], S.cl_len[sym]);
                        if (sym == 16) w.put(S.cl_extra[q], 2);
                        if (sym == 17) w.put(S.cl_extra[q], 3);
                        if (sym == 18) w.put(S.cl_extra[q], 7);
                    }
                    w.finish();
                }
                if (s0 < bl) {
                    BitWriter w(slot, hb + off);
                    for (int q = 0; q < ntok; q++) {
                        const uint32_t tk = tok[q * DF_T];
                        if (tk & TOK_MATCH) {
                            uint32_t ls, le, lx, ds, de, dx;
                            len_code((tk >> 16) & 0x1FF, ls, le, lx);
                            dist_code(tk & 0xFFFF ? tk & 0xFFFF : 65536, ds, de, dx);
                            w.put(S.lit_code[ls], S.lit_len[ls]);
                            if (le) w.put(lx, (int)le);
                            w.put(S.d_code[ds], S.d_len[ds]);
                            if (de) w.put(dx, (int)de);
                        } else {
                            w.put(S.lit_code[tk], S.lit_len[tk]);
                        }
                    }
                    if (s1 == bl) w.put(S.lit_code[256], S.lit_len[256]);
                    w.finish();
                }
            } else {
                nbytes = 0xFFFFFFFFu;
            }
        }
        __syncthreads();
        // 9. sizes and CRC (the body is already in the block's slot)
        if (t == 0) {
            body_len[b] = nbytes == 0xFFFFFFFFu ? STORED : nbytes;
            crc_out[b] = crc;
            member_bytes[b] = 18 + (nbytes == 0xFFFFFFFFu ? (uint32_t)bl + 5 : nbytes) + 8;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t member_byte(int64_t m, int64_t nb, int64_t n, uint32_t j, uint32_t size,
                                                const uint8_t *__restrict__ src,
                                                const uint8_t *__restrict__ slot, uint32_t body,
                                                uint32_t crc) {
    if (m == nb) {  // EOF member: empty block, CRC 0, ISIZE 0
        constexpr uint8_t eof[28] = {0x1f, 0x8b, 0x08, 0x04, 0, 0, 0, 0, 0, 0xff, 0x06, 0, 0x42, 0x43,
                                     0x02, 0, 0x1b, 0, 0x03, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < 28; k++) v = (uint32_t)k == j ? eof[k] : v;
        return v;
    }
    const int64_t base = m * DF_IN;
    const uint32_t bl = (uint32_t)(n - base < DF_IN ? n - base : DF_IN);
    if (j < 18) {
        if (j == 16) return (size - 1) & 255;
        if (j == 17) return (size - 1) >> 8;
        constexpr uint32_t h0 = 0x04088b1fu, h2 = 0x0006ff00u, h3 = 0x00024342u;  // bytes 0-3, 8-11, 12-15
        return j < 4 ? (h0 >> (8 * j)) & 255 : j < 8 ? 0u : j < 12 ? (h2 >> (8 * (j - 8))) & 255 : (h3 >> (8 * (j - 12))) & 255;
    }
    if (j < size - 8) {
        const uint32_t p = j - 18;
        if (body != STORED) return slot[p];
        if (p >= 5) return src[base + p - 5];
        const uint32_t nl = ~bl & 0xFFFF;
        return p == 0 ? 1u : p == 1 ? bl & 255 : p == 2 ? bl >> 8 : p == 3 ? nl & 255 : nl >> 8;
    }
    const uint32_t q = j - (size - 8);
    return q < 4 ? (crc >> (8 * q)) & 255 : (bl >> (8 * (q - 4))) & 255;
}

// One wave per member; lanes build 4-byte-aligned words of the output.
__global__ __launch_bounds__(256) void bgzf_assemble_kernel(const uint8_t *__restrict__ src, int64_t n, int64_t nb,
                                                            int64_t n_members,
                                                            const uint8_t *__restrict__ slots,
                                                            const uint32_t *__restrict__ body_len,
                                                            const uint32_t *__restrict__ crc_in,
                                                            const uint32_t *__restrict__ member_bytes,
                                                            const uint64_t *__restrict__ member_off,
                                                            uint8_t *__restrict__ dst, int64_t dst_cap) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)__builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    uint32_t *const dst32 = reinterpret_cast<uint32_t *>(dst);
    for (int64_t m = wave; m < n_members; m += (int64_t)gridDim.x * 4) {
        const uint32_t size = member_bytes[m];
        const int64_t off = (int64_t)member_off[m];
        if (off + size > dst_cap) continue;
        const uint32_t body = m < nb ? body_len[m] : 0;
        const uint32_t crc = m < nb ? crc_in[m] : 0;
        const uint8_t *slot = slots + m * DF_SLOT;
        const int64_t d0 = off >> 2, d1 = (off + size - 1) >> 2;
        for (int64_t d = d0 + lane; d <= d1; d += 64) {
            const int64_t b0 = 4 * d - off;
            if (b0 >= 0 && b0 + 4 <= size) {
                uint32_t w = 0;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    w |= member_byte(m, nb, n, (uint32_t)b0 + k, size, src, slot, body, crc) << (8 * k);
                dst32[d] = w;
            } else {
                for (int k = 0; k < 4; k++) {
                    const int64_t j = b0 + k;
                    if (j >= 0 && j < size)
                        dst[off + j] = (uint8_t)member_byte(m, nb, n, (uint32_t)j, size, src, slot, body, crc);
                }
            }
        }
    }
}

__global__ void set_u32_kernel(uint32_t *p, uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

constexpr int DF_MAX_GRID = 256;  // resident workgroups at most (one per CU)

struct DfScratch {
    size_t slots, body, crc, bytes, off, partial, cand, tok, total;
};

DfScratch df_layout(int64_t n) {
    const int64_t nb = (n + DF_IN - 1) / DF_IN, nm = nb + 1;
    const int64_t g = nb < DF_MAX_GRID ? (nb > 0 ? nb : 1) : DF_MAX_GRID;
    DfScratch s;
    size_t o = 0;
    auto take = [&](size_t b) {
        const size_t r = o;
        o += (b + 255) & ~(size_t)255;
        return r;
    };
    s.slots = take((size_t)(nb > 0 ? nb : 1) * DF_SLOT);
    s.body = take((size_t)nm * 4);
    s.crc = take((size_t)nm * 4);
    s.bytes = take((size_t)nm * 4);
    s.off = take((size_t)(nm + 1) * 8);
    s.partial = take(scan_scratch_bytes(nm));
    s.cand = take((size_t)g * DF_IN * 2);
    s.tok = take((size_t)g * DF_T * DF_SEG * 4);
    s.total = o;
    return s;
}

}  // namespace

extern "C" {

int64_t ma_bgzf_dev_scratch_bytes(int64_t n) { return (int64_t)df_layout(n < 0 ? 0 : n).total; }

int ma_bgzf_compress_dev(void *stream, const uint8_t *d_src, int64_t n, int32_t level, int32_t add_eof,
                         uint8_t *d_dst, int64_t dst_cap, uint64_t *d_total, void *d_scratch,
                         int64_t scratch_bytes) {
    MA_REQUIRE(n >= 0 && (n == 0 || d_src) && d_dst && d_total, "ma_bgzf_compress_dev: bad arguments");
    MA_REQUIRE(level >= 0 && level <= 9, "ma_bgzf_compress_dev: level outside 0..9");
    MA_REQUIRE(((uintptr_t)d_src & 3) == 0 && ((uintptr_t)d_dst & 3) == 0,
               "ma_bgzf_compress_dev: buffers not 4-byte aligned");
    const DfScratch L = df_layout(n);
    MA_REQUIRE(scratch_bytes >= (int64_t)L.total, "ma_bgzf_compress_dev: scratch too small");
    hipStream_t st = (hipStream_t)stream;
    char *s = (char *)d_scratch;
    const int64_t nb = (n + DF_IN - 1) / DF_IN;
    const int64_t nm = nb + (add_eof ? 1 : 0);
    uint32_t *body = (uint32_t *)(s + L.body), *crc = (uint32_t *)(s + L.crc), *bytes = (uint32_t *)(s + L.bytes);
    uint64_t *off = (uint64_t *)(s + L.off);
    if (nb > 0) {
        int dev = 0, n_cu = 256;
        MA_HIP(hipGetDevice(&dev));
        MA_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
        int64_t grid = nb < n_cu ? nb : n_cu;
        if (grid > DF_MAX_GRID) grid = DF_MAX_GRID;
        bgzf_deflate_kernel<<<(unsigned)grid, DF_T, 0, st>>>(d_src, n, nb, level, (uint32_t *)(s + L.slots),
                                                              body, crc, bytes, (uint16_t *)(s + L.cand),
                                                              (uint32_t *)(s + L.tok));
        MA_LAUNCHED();
    }
    if (add_eof) {
        set_u32_kernel<<<1, 64, 0, st>>>(bytes + nb, 28u);
        MA_LAUNCHED();
    }
    if (nm == 0) return fill_u64(st, d_total, 1, 0);
    MA_TRY(exclusive_scan_u32(st, nm, bytes, off, (uint64_t *)(s + L.partial), d_total));
    const int64_t blocks = (nm + 3) / 4 < 65536 ? (nm + 3) / 4 : 65536;
    bgzf_assemble_kernel<<<(unsigned)blocks, 256, 0, st>>>(d_src, n, nb, nm, (const uint8_t *)(s + L.slots), body,
                                                           crc, bytes, off, d_dst, dst_cap);
    MA_LAUNCHED();
    return MA_OK;
}

}  // extern "C"
