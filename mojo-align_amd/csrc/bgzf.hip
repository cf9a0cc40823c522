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
//     in LDS;
//   * LZ77 by wave: each of the 16 waves parses its own 4096-byte sub-block
//     with its own 1024-entry hash table in LDS, 64 positions (a segment) per
//     round: a position's candidate is the latest position with its 4-byte
//     hash from an earlier round (the table is seeded with the previous
//     sub-block's last 512 positions), compared 4 bytes at a time from windows
//     read once per round; the distance-1 run's length comes from a ballot of
//     "byte equals the one before"; matches of 4..64 bytes stay inside the
//     segment, which is parsed greedily in scalar code over the lanes' match
//     masks; a 2-byte token per position goes to a scratch, and each wave
//     counts its symbols (two 16-bit counters per LDS word);
//   * one dynamic-Huffman block (RFC 1951 3.2.7) per member, every code's
//     lengths optimal under its limit by package-merge (Larmore and
//     Hirschberg): literal/length and distance codes (15 bits) at once across
//     the workgroup, one barrier per level, over symbols ranked by packed
//     (count, symbol) keys; the code-length code (7 bits) in one wave; the
//     code-length sequence's runs from a ballot of run starts; canonical codes
//     in parallel; a stored block when that is not smaller;
//   * each wave's body bits follow from its symbol counts and the code
//     lengths, so the sub-blocks' bit offsets need no second pass; the body is
//     assembled in LDS (where the hash tables were) by OR: the header's fields
//     and code-length symbols at offsets from a block scan, every token's bits
//     at offsets from a DPP wave scan (each lane writes 8 consecutive tokens);
//     then it is copied to the block's output slot;
//   * CRC32: per-thread CRCs of 64-byte segments (slicing by 4) joined by the
//     GF(2) shift crc(A||B) = crc(A) * x^(8|B|) + crc(B) mod P;
//   * a second kernel writes every member (18-byte header, body, trailer) at
//     its offset from a scan of the member sizes, and the EOF member.
#include "ma_sort.hpp"

using namespace ma;

namespace {

constexpr int DF_T = 1024;            // threads per workgroup
constexpr int DF_SEG = 64;            // input bytes per thread's CRC segment
constexpr int DF_IN = 65280;          // bytes per BGZF block
constexpr int DF_OUT_WORDS = 16384;   // 64 KiB of deflate body per block at most
constexpr int64_t DF_SLOT = 65536;    // staging bytes per block
constexpr int NLIT = 286;             // literal/length alphabet
constexpr int NDC = 30;               // distance alphabet
constexpr int NCL = 19;               // code-length alphabet
constexpr int NSEQ = NLIT + NDC;      // code-length sequence entries at most
static_assert(NSEQ <= DF_T, "one code-length symbol per thread");
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

constexpr int DF_W = DF_T / 64;      // waves per workgroup
constexpr int SUB = 4096;            // bytes of the block parsed by one wave (64 segments)
constexpr int PRE = 512;             // positions before a sub-block seeded into its wave's table
constexpr int TAB_BITS = 10;         // per-wave match-candidate table: 1024 entries
constexpr int NSYM = NLIT + NDC;     // literal/length and distance symbols, counted together
constexpr int HW = NSYM / 2;         // per-wave counters: two 16-bit counts per word
constexpr int PM_L = 15;             // code length limit of the literal/length and distance codes
constexpr int PM_D = 640;            // first thread of the distance code's package-merge
constexpr int CL_L = 7;              // code length limit of the code-length code

struct DfShared {
    uint32_t data[1 + DF_IN / 4 + 24];  // a guard word, the block, padding: every window read stays inside
    union alignas(16) {
        uint32_t tab[DF_W][1 << TAB_BITS];  // per wave: 1 + the latest earlier-round position per hash
        struct {                            // then the package-merge lists of both codes
            uint32_t a[2][2 * NLIT];        // weights of a list in merged order (two levels)
            uint32_t da[2][2 * NDC];
            uint16_t lc[PM_L][2 * NLIT + 1];  // per level and place p: leaves among the first p items
            uint16_t dlc[PM_L][2 * NDC + 1];
            uint32_t x[2][PM_L + 1];        // per level: leaves among the selected items
        } pm;
        uint32_t out[DF_OUT_WORDS];         // then the deflate body, assembled by OR
    } u;
    uint32_t hist[DF_W][HW];        // per wave: symbol counts of its sub-block (16-bit halves)
    uint32_t crc4[4][256];          // slicing-by-4 CRC tables
    uint32_t xpow[40];              // x^(2^k) mod P (reflected), k = 0..39
    uint32_t crcw[DF_W];            // per-wave XOR of the shifted segment CRCs
    uint32_t wextra[DF_W];          // per-wave extra bits of lengths and distances
    uint32_t wbits[DF_W];           // per-wave body bits
    uint32_t sf[NLIT];              // frequencies in ranked order
    uint32_t dsf[NDC];
    uint32_t clblc[32];
    uint32_t wsum[DF_T / 64];
    uint32_t freq[NLIT];
    uint32_t dfreq[NDC];
    alignas(4) uint8_t lit_len[NLIT];
    uint32_t lpk[NLIT];             // bit-reversed canonical code | its length << 16
    uint32_t mpk[65];               // per match length: its length code + extra bits | their count << 24
    alignas(4) uint8_t d_len[NDC];  // (read four at a time)
    uint32_t dpk[NDC];              // as lpk, for the distance code
    uint16_t sorted[NLIT];          // used symbols by ascending (freq, symbol)
    uint32_t rnk[NLIT];             // a literal/length symbol's rank
    uint32_t key[NLIT];             // count << 9 | symbol, all ones for an unused symbol
    uint64_t rstart[(NSEQ + 63) / 64];  // code-length sequence: where runs of equal lengths start
    uint16_t dsorted[NDC];
    uint32_t dblc[32];
    uint8_t cl_sym[NSEQ];
    uint8_t cl_extra[NSEQ];
    uint32_t cl_freq[NCL];
    alignas(4) uint8_t cl_len[NCL];
    alignas(8) uint32_t cla[2][2 * NCL];    // the code-length code's package-merge lists
    uint16_t cllc[CL_L][2 * NCL + 1];
    uint32_t clx[CL_L];
    uint16_t cl_code[NCL];
    uint16_t cl_sorted[NCL];
    uint32_t blc[32];
    uint32_t misc[8];               // 0 used lit/len symbols, 2 last used lit/len symbol + 1,
                                    // 3 last sent code-length code length + 1, 5 used distance
                                    // symbols, 6 last used distance symbol + 1, 7 used CL symbols
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

// The 4 bytes at byte offset x of a word array (two aligned reads).
__device__ __forceinline__ uint32_t load4(const uint32_t *dw, int x) {
    const int a = x >> 2;
    return __builtin_amdgcn_alignbyte(dw[a + 1], dw[a], (uint32_t)(x & 3));
}

__device__ __forceinline__ uint32_t hash_at(const uint32_t *dw, int p) {
    return (load4(dw, p) * 2654435761u) >> (32 - TAB_BITS);
}

// Length (at most 32) of the common prefix of two byte strings given as
// eight 4-byte windows each.
__device__ __forceinline__ int prefix32(const uint32_t (&a)[8], const uint32_t (&b)[8]) {
    int l = 32;
#pragma unroll
    for (int j = 7; j >= 0; j--) {
        const uint32_t x = a[j] ^ b[j];
        if (x) l = 4 * j + (int)(__builtin_ctz(x) >> 3);
    }
    return l;
}

// The eight 4-byte windows of the 32 bytes at byte offset x (nine aligned reads
// issued together).
__device__ __forceinline__ void windows32(const uint32_t *dw, int x, uint32_t (&win)[8]) {
    const int a = x >> 2;
    const uint32_t s = (uint32_t)(x & 3);
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = dw[a + j];
#pragma unroll
    for (int j = 0; j < 8; j++) win[j] = __builtin_amdgcn_alignbyte(w[j + 1], w[j], s);
}

// Common prefix of the bytes at q and p (q < p) from byte 32 on, at most 32 more.
__device__ __forceinline__ int match_more(const uint32_t *dw, int q, int p) {
    uint32_t a[8], b[8];
    windows32(dw, q + 32, a);
    windows32(dw, p + 32, b);
    return prefix32(a, b);
}

// ORs the n low bits of v into the bit stream at bit offset o (LSB first).
__device__ __forceinline__ void or_bits(uint32_t *out, uint32_t o, uint64_t v, int n) {
    if (n == 0) return;
    const uint32_t wi = o >> 5, s = o & 31;
    const uint64_t lo = v << s;
    const uint32_t hi = s ? (uint32_t)(v >> (64 - s)) : 0u;
    if ((uint32_t)lo) atomicOr(&out[wi], (uint32_t)lo);
    if ((uint32_t)(lo >> 32)) atomicOr(&out[wi + 1], (uint32_t)(lo >> 32));
    if (hi) atomicOr(&out[wi + 2], hi);
}

// Bits [a, b) of a 64-bit mask, 0 <= a <= b <= 64, a < 64.
__device__ __forceinline__ uint64_t span64(int a, int b) {
    const uint64_t hi = b >= 64 ? ~0ull : (1ull << b) - 1;
    return hi & ~((1ull << a) - 1);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Orders a wave's LDS writes before its later reads by other lanes.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive wave scan by DPP: shifts of 1, 2, 4, 8 lanes inside each row of 16,
// then row 0's total into row 1 and row 2's into row 3 (row_bcast:15), then the
// first two rows' total into rows 2 and 3 (row_bcast:31); lanes whose source is
// outside their row, and rows outside the row mask, add 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
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

// Canonical code of symbol s (bit-reversed): the first code of its length
// plus the number of smaller symbols of that length.
__device__ __forceinline__ uint16_t canon_code(const uint8_t *len, const uint32_t *blc, int s) {
    const int l = len[s];
    if (l == 0) return 0;
    uint32_t code = 0;
    for (int k = 1; k <= l; k++) code = (code + (k > 1 ? blc[k - 1] : 0)) << 1;
    // symbols before s with length l, four length bytes per step (zero-byte count of len ^ l)
    const uint32_t *lw = reinterpret_cast<const uint32_t *>(len);
    const uint32_t pat = (uint32_t)l * 0x01010101u;
    int u = 0;
    for (; u + 4 <= s; u += 4) {
        const uint32_t x = lw[u >> 2] ^ pat;
        code += (uint32_t)__popc(~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu));
    }
    for (; u < s; u++) code += len[u] == l;
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

// First index in [0, n) of an ascending array whose value is >= v (strict: > v).
template <bool STRICT, typename F>
__device__ __forceinline__ int search_first(int n, uint32_t v, F at) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint32_t x = at(mid);
        if (STRICT ? x <= v : x < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// One level of package-merge for a code of n leaves with ascending weights w[]:
// list A (size s, ascending) becomes the next level's list B = merge(leaves,
// packages of consecutive pairs of A), a leaf before a package of equal weight;
// lc[p] gets the number of leaves among B's first p items (p <= |B|). Thread
// e < n + s / 2 places element e.
__device__ __forceinline__ void pm_level(int e, int n, int s, const uint32_t *w, const uint32_t *A, uint32_t *B,
                                         uint16_t *lc) {
    const int np = s >> 1;
    if (e >= n + np) return;
    const uint2 *P = reinterpret_cast<const uint2 *>(A);
    if (e == 0) lc[n + np] = (uint16_t)n;
    if (e < n) {
        const uint32_t v = w[e];
        const int at = e + search_first<false>(np, v, [&](int k) { const uint2 q = P[k]; return q.x + q.y; });
        B[at] = v;
        lc[at] = (uint16_t)e;
    } else {
        const int k = e - n;
        const uint2 q = P[k];
        const uint32_t v = q.x + q.y;
        const int c = search_first<true>(n, v, [&](int i) { return w[i]; });
        B[k + c] = v;
        lc[k + c] = (uint16_t)c;
    }
}

// Leaves x_l among the first c_l items of each level's list, level 1 first:
// c_1 = 2n - 2 items, and the packages among a level's first c_l items are the
// first c_(l+1) = 2 (c_l - x_l) items of the level below. One thread.
template <int W>
__device__ __forceinline__ void pm_select(int n, const uint16_t (*lc)[W], uint32_t *x) {
    int c = 2 * n - 2;
    for (int l = 0; l < PM_L; l++) {
        const int xl = lc[l][c];
        x[l] = (uint32_t)xl;
        c = 2 * (c - xl);
    }
}

// Tokens, two bytes per block position: a literal byte b is TOK_LIT | b, a
// match starting there TOK_MATCH | distance (at most SUB + PRE), and the
// position after a match start (always inside the match) holds its length;
// every other position inside a match is 0.
constexpr uint32_t TOK_LIT = 1u << 15;
constexpr uint32_t TOK_MATCH = 1u << 14;
static_assert(SUB + PRE < (int)TOK_MATCH, "a distance fits below the match flag");
constexpr int MIN_MATCH = 4;

__device__ __forceinline__ void count_sym(uint32_t *hist, uint32_t u) {
    atomicAdd(&hist[u >> 1], 1u << (16 * (u & 1)));
}

__global__ __launch_bounds__(DF_T) void bgzf_deflate_kernel(const uint8_t *__restrict__ src, int64_t n,
                                                            int64_t nb, int level,
                                                            uint32_t *__restrict__ slots,
                                                            uint32_t *__restrict__ body_len,
                                                            uint32_t *__restrict__ crc_out,
                                                            uint32_t *__restrict__ member_bytes,
                                                            uint16_t *__restrict__ tok_all) {
    __shared__ DfShared S;
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);  // uniform: sub-block bounds stay scalar
    uint32_t *const dw = S.data + 1;  // the block's words (dw[-1] is the guard)
    const uint8_t *const db = reinterpret_cast<const uint8_t *>(dw);
    uint16_t *const tok = tok_all + (int64_t)blockIdx.x * DF_IN;  // this workgroup's tokens
    if (t < 256) {
        uint32_t c = (uint32_t)t;
        for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ CRC_POLY : c >> 1;
        S.crc4[0][t] = c;
    }
    if (t == 0) {
        uint32_t x = 0x40000000u;  // x^1
        for (int k = 0; k < 40; k++) {
            S.xpow[k] = x;
            x = multmodp(x, x);
        }
    }
    __syncthreads();
    if (t < 256) {  // slicing by 4 (Kounavis and Berry): table k advances a byte k more places
        uint32_t c = S.crc4[0][t];
        for (int k = 1; k < 4; k++) {
            c = (c >> 8) ^ S.crc4[0][c & 255];
            S.crc4[k][t] = c;
        }
    }
    // the CRC shift over the bytes after this thread's segment in a full block
    const uint32_t sh_full = t < DF_IN / DF_SEG ? x8n(S.xpow, (uint32_t)(DF_IN - DF_SEG * (t + 1))) : 0u;
    __syncthreads();
    for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t base = b * DF_IN;
        const int bl = (int)(n - base < DF_IN ? n - base : DF_IN);
        uint32_t *const slot = slots + b * (DF_SLOT / 4);
        // 1. stage the block: four words' loads issued before their stores
        {
            const uint8_t *blk = src + base;
            for (int i0 = 4 * t; i0 < bl; i0 += 4 * 4 * DF_T) {
                uint32_t v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int i = i0 + 4 * DF_T * k;
                    v[k] = i + 4 <= bl ? *reinterpret_cast<const uint32_t *>(blk + i) : 0u;
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int i = i0 + 4 * DF_T * k;
                    if (i < bl) {
                        uint32_t x = v[k];
                        if (i + 4 > bl)
                            for (int j = 0; j < bl - i; j++) x |= (uint32_t)blk[i + j] << (8 * j);
                        dw[i >> 2] = x;
                    }
                }
            }
        }
        __syncthreads();
        // 2. CRC of this thread's 64-byte segment (slicing by 4), shifted to its place
        {
            const int s0 = t * DF_SEG, s1 = min(bl, s0 + DF_SEG);
            uint32_t c = 0xFFFFFFFFu, v = 0;
            if (s0 < bl) {
                int i = s0;
                for (; i + 4 <= s1; i += 4) {
                    c ^= dw[i >> 2];
                    c = S.crc4[3][c & 255] ^ S.crc4[2][(c >> 8) & 255] ^ S.crc4[1][(c >> 16) & 255] ^
                        S.crc4[0][c >> 24];
                }
                for (; i < s1; i++) c = S.crc4[0][(c ^ db[i]) & 255] ^ (c >> 8);
                // crc(S_0 .. S_k) = XOR over segments of crc(S_t) * x^(8 * bytes after S_t) mod P
                v = multmodp(bl == DF_IN ? sh_full : x8n(S.xpow, (uint32_t)(bl - s1)), ~c);
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v ^= __shfl_xor(v, o, 64);
            if (lane == 0) S.crcw[w] = v;
        }
        // 3. LZ77, one wave per 4096-byte sub-block, 64 positions (a segment) per round:
        //    a position's candidate is the latest position of an earlier round (or of the
        //    previous sub-block's last PRE bytes) with its 4-byte hash; the longer of its
        //    match and the distance-1 run (4..64 bytes, inside the segment) is its match;
        //    the segment is parsed greedily from those, in scalar code over lane masks
        const int sb0 = w * SUB, sb1 = min(sb0 + SUB, bl);
        if (level > 0) {
            uint32_t *const tab = S.u.tab[w];
            uint32_t *const hist = S.hist[w];
            for (int k = lane; k < (1 << TAB_BITS); k += 64) tab[k] = 0;
            for (int k = lane; k < HW; k += 64) hist[k] = 0;
            uint32_t extra = 0;
            if (sb0 < sb1) {
                for (int p = (sb0 >= PRE ? sb0 - PRE : 0) + lane; p < sb0; p += 64)
                    if (p + 4 <= bl) atomicMax(&tab[hash_at(dw, p)], (uint32_t)(p + 1));
                for (int r0 = sb0; r0 < sb1; r0 += 64) {
                    const int P = r0 + lane;
                    const int E = min(r0 + 64, sb1);
                    const bool valid = P < E;
                    const bool ins = valid && P + 4 <= bl;
                    // the 36 bytes from P - 4 (aligned down), read once (the guard word and the
                    // padding keep every lane's reads inside): P's windows, the hash, byte P - 1
                    const int pa = P >> 2;
                    const uint32_t ps = (uint32_t)(P & 3);
                    uint32_t pw[10];
#pragma unroll
                    for (int j = 0; j < 10; j++) pw[j] = dw[pa - 1 + j];
                    uint32_t pwin[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) pwin[j] = __builtin_amdgcn_alignbyte(pw[j + 2], pw[j + 1], ps);
                    // bit k: byte r0 + k equals the byte before it (a distance-1 run goes on)
                    const uint32_t prev = ps ? (pw[1] >> (8 * (ps - 1))) & 255u : pw[0] >> 24;
                    const uint64_t eqp = __ballot(valid && (pwin[0] & 255u) == prev);
                    uint32_t h = 0, c = 0;
                    if (ins) {
                        h = (pwin[0] * 2654435761u) >> (32 - TAB_BITS);
                        c = tab[h];  // the wave's reads complete before its updates below
                    }
                    if (ins) atomicMax(&tab[h], (uint32_t)(P + 1));
                    int best = 0, bd = 0;
                    const int maxl = E - P;
                    if (valid && maxl >= MIN_MATCH) {
                        if (c) {
                            const int q = (int)c - 1;
                            uint32_t qwin[8];
                            windows32(dw, q, qwin);
                            int l = prefix32(qwin, pwin);
                            if (l == 32 && maxl > 32) l += match_more(dw, q, P);
                            l = min(l, maxl);
                            best = l;
                            bd = P - q;
                        }
                        if (P >= 1 && best < maxl) {  // the run: the set bits of eqp from P on
                            const uint64_t stop = ~eqp >> lane;
                            const int l = min(stop ? (int)__builtin_ctzll(stop) : 64 - lane, maxl);
                            if (l > best) {
                                best = l;
                                bd = 1;
                            }
                        }
                    }
                    const uint64_t M = __ballot(best >= MIN_MATCH);
                    const int nv = E - r0;
                    uint64_t lit = 0, st = 0;
                    int p = 0;
                    while (p < nv) {
                        const uint64_t rest = M >> p << p;
                        if (!rest) {
                            lit |= span64(p, nv);
                            break;
                        }
                        const int q = __builtin_ctzll(rest);
                        lit |= span64(p, q);
                        st |= 1ull << q;
                        p = q + __builtin_amdgcn_readlane(best, q);
                    }
                    // a covered position right after a match start carries the match's length
                    const int prev_best = __shfl_up(best, 1, 64);
                    uint32_t tk = lane >= 1 && ((st >> (lane - 1)) & 1) ? (uint32_t)prev_best : 0u;
                    if ((lit >> lane) & 1) {
                        const uint32_t by = db[P];
                        tk = TOK_LIT | by;
                        count_sym(hist, by);
                    } else if ((st >> lane) & 1) {
                        uint32_t ls, le, lx, ds, de, dx;
                        len_code((uint32_t)best, ls, le, lx);
                        dist_code((uint32_t)bd, ds, de, dx);
                        tk = TOK_MATCH | (uint32_t)bd;
                        count_sym(hist, ls);
                        count_sym(hist, NLIT + ds);
                        extra += le + de;
                    }
                    if (valid) tok[P] = (uint16_t)tk;
                }
            }
            extra = wave_sum(extra);
            if (lane == 0) S.wextra[w] = extra;
        }
        __syncthreads();
        const uint32_t crc = [&] {
            uint32_t x = 0;
            for (int v = 0; v < DF_W; v++) x ^= S.crcw[v];
            return x;
        }();
        uint32_t nbytes = 0xFFFFFFFFu;
        if (level > 0) {
            // 4. block symbol counts; the output words cleared (the tables are done)
            if (t < NSYM) {
                uint32_t f = 0;
                for (int v = 0; v < DF_W; v++) f += (S.hist[v][t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
                if (t < NLIT) {
                    f += t == 256 ? 1u : 0u;  // end of block
                    S.freq[t] = f;
                    S.key[t] = f ? f << 9 | (uint32_t)t : 0xFFFFFFFFu;
                    S.lit_len[t] = 0;
                    S.rnk[t] = 0;
                } else {
                    S.dfreq[t - NLIT] = f;
                    S.d_len[t - NLIT] = 0;
                }
            }
            if (t < NCL) {
                S.cl_freq[t] = 0;
                S.cl_len[t] = 0;
            }
            if (t < 8) S.misc[t] = 0;
            __syncthreads();
            // 5. the symbols of both codes ranked by (count, symbol) in parallel
            //    (a literal/length symbol's rank summed over three thirds of the alphabet)
            if (t < 3 * NLIT) {
                const int part = t / NLIT, sy = t - part * NLIT;
                const uint32_t ks = S.key[sy];
                if (ks != 0xFFFFFFFFu) {
                    uint32_t r = 0;
                    const int u1 = min(NLIT, part * 96 + 96);
#pragma unroll 8
                    for (int u = part * 96; u < u1; u++) r += S.key[u] < ks ? 1u : 0u;
                    atomicAdd(&S.rnk[sy], r);
                }
            }
            rank_symbols(S.dfreq, NDC, S.dsorted, S.dsf, &S.misc[5], 896);
            __syncthreads();
            if (t < NLIT && S.freq[t]) {
                const uint32_t r = S.rnk[t];
                S.sorted[r] = (uint16_t)t;
                S.sf[r] = S.freq[t];
                atomicAdd(&S.misc[0], 1u);
            }
            if (t < 32) {
                S.blc[t] = 0;
                S.dblc[t] = 0;
                S.clblc[t] = 0;
            }
            __syncthreads();
            // 6. code lengths limited to 15 bits, optimal for the counts: package-merge
            //    (Larmore and Hirschberg) over the ranked weights, both codes at once
            //    (threads from 0 and from PM_D), one barrier per level
            const int m = (int)S.misc[0], md = (int)S.misc[5];  // m >= 2: a byte and end of block
            const bool dtree = md >= 2;
            if (t <= m) {  // the deepest level: the leaves alone
                if (t < m) S.u.pm.a[1][t] = S.sf[t];
                S.u.pm.lc[PM_L - 1][t] = (uint16_t)t;
            }
            if (dtree && t >= PM_D && t - PM_D <= md) {
                if (t - PM_D < md) S.u.pm.da[1][t - PM_D] = S.dsf[t - PM_D];
                S.u.pm.dlc[PM_L - 1][t - PM_D] = (uint16_t)(t - PM_D);
            }
            __syncthreads();
            {
                int s = m, sd = md, cur = 1;
#pragma unroll 1
                for (int l = PM_L - 2; l >= 0; l--) {  // list of level l + 1 from that of level l + 2
                    pm_level(t, m, s, S.sf, S.u.pm.a[cur], S.u.pm.a[cur ^ 1], S.u.pm.lc[l]);
                    if (dtree && t >= PM_D)
                        pm_level(t - PM_D, md, sd, S.dsf, S.u.pm.da[cur], S.u.pm.da[cur ^ 1], S.u.pm.dlc[l]);
                    __syncthreads();
                    s = m + (s >> 1);
                    sd = md + (sd >> 1);
                    cur ^= 1;
                }
            }
            if (t == 0) pm_select(m, S.u.pm.lc, S.u.pm.x[0]);
            if (t == 64 && dtree) pm_select(md, S.u.pm.dlc, S.u.pm.x[1]);
            if (t == 128 && !dtree) {  // no match: two 1-bit codes; one distance symbol: it and a spare
                if (md == 0) {
                    S.d_len[0] = S.d_len[1] = 1;
                } else {
                    const int s = S.dsorted[0];
                    S.d_len[s] = 1;
                    S.d_len[s == 0 ? 1 : 0] = 1;
                }
                S.dblc[1] = 2;
            }
            __syncthreads();
            if (t < m) {  // a leaf's length: the levels whose selection holds it
                int len = 0;
#pragma unroll 1
                for (int l = 0; l < PM_L; l++) len += t < (int)S.u.pm.x[0][l] ? 1 : 0;
                S.lit_len[S.sorted[t]] = (uint8_t)len;
                atomicAdd(&S.blc[len], 1u);
            }
            if (dtree && t >= PM_D && t - PM_D < md) {
                int len = 0;
#pragma unroll 1
                for (int l = 0; l < PM_L; l++) len += t - PM_D < (int)S.u.pm.x[1][l] ? 1 : 0;
                S.d_len[S.dsorted[t - PM_D]] = (uint8_t)len;
                atomicAdd(&S.dblc[len], 1u);
            }
            __syncthreads();
            if (t < NLIT) {
                S.lpk[t] = canon_code(S.lit_len, S.blc, t) | (uint32_t)S.lit_len[t] << 16;
                if (S.lit_len[t]) atomicMax(&S.misc[2], (uint32_t)(t + 1));
            }
            if (t >= 512 && t - 512 < NDC) {
                S.dpk[t - 512] = canon_code(S.d_len, S.dblc, t - 512) | (uint32_t)S.d_len[t - 512] << 16;
                if (S.d_len[t - 512]) atomicMax(&S.misc[6], (uint32_t)(t - 511));
            }
            for (int k = t; k < DF_OUT_WORDS; k += DF_T) S.u.out[k] = 0;  // the lists are done
            {  // this wave's body bits: its symbol counts times their code lengths, plus extra bits
                uint32_t bits = 0;
                for (int u = lane; u < NSYM; u += 64) {
                    const uint32_t cnt = (S.hist[w][u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
                    bits += cnt * (u < NLIT ? S.lit_len[u] : S.d_len[u - NLIT]);
                }
                bits = wave_sum(bits);
                if (lane == 0) S.wbits[w] = bits + S.wextra[w];
            }
            __syncthreads();
            if (t >= DF_T - 64 && t - (DF_T - 64) + MIN_MATCH <= 64) {  // the match lengths' codes
                const uint32_t L = (uint32_t)(t - (DF_T - 64) + MIN_MATCH);
                uint32_t ls, le, lx;
                len_code(L, ls, le, lx);
                const uint32_t lk = S.lpk[ls], ll = lk >> 16;
                S.mpk[L] = ((lk & 0xFFFFu) | lx << ll) | (ll + le) << 24;
            }
            // code-length sequence (literal/length then distance lengths): maximal runs
            // found in parallel, each coded at its scanned offset
            const int hlit = max(257, (int)S.misc[2]), hdist = max(1, (int)S.misc[6]);
            const int total = hlit + hdist;
            int rv = 0, run = 0;
            uint32_t cnt = 0;
            bool start = false;
            if (t < total) {
                rv = t < hlit ? S.lit_len[t] : S.d_len[t - hlit];
                const int pv = t == 0 ? -1 : (t - 1 < hlit ? S.lit_len[t - 1] : S.d_len[t - 1 - hlit]);
                start = rv != pv;
            }
            {
                const uint64_t sm = __ballot(start);
                if (w < (NSEQ + 63) / 64 && lane == 0) S.rstart[w] = sm;
            }
            __syncthreads();
            if (start) {  // a run ends where the next one starts
                int wi = t >> 6;
                uint64_t rest = S.rstart[wi] & ~((2ull << (t & 63)) - 1);
                while (!rest && ++wi < (NSEQ + 63) / 64) rest = S.rstart[wi];
                const int next = rest ? 64 * wi + (int)__builtin_ctzll(rest) : total;
                run = min(next, total) - t;
                cnt = (uint32_t)rle_emit(rv, run, nullptr, nullptr);
            }
            uint32_t ns;
            const uint32_t co = block_excl_scan_1024(cnt, S.wsum, ns);
            if (cnt) rle_emit(rv, run, S.cl_sym + co, S.cl_extra + co);
            __syncthreads();
            for (uint32_t q = (uint32_t)t; q < ns; q += DF_T) atomicAdd(&S.cl_freq[S.cl_sym[q]], 1u);
            __syncthreads();
            rank_symbols(S.cl_freq, NCL, S.cl_sorted, S.sf, &S.misc[7], 0);
            __syncthreads();
            if (w == 0) {  // the code-length code: package-merge limited to 7 bits, one wave
                const int nc = (int)S.misc[7];
                if (nc == 1) {  // one symbol: it and a spare, both 1 bit
                    if (lane == 0) {
                        const int sy = S.cl_sorted[0];
                        S.cl_len[sy] = 1;
                        S.cl_len[sy == 0 ? 1 : 0] = 1;
                        S.clblc[1] = 2;
                    }
                } else {
                    if (lane <= nc) {
                        if (lane < nc) S.cla[1][lane] = S.sf[lane];
                        S.cllc[CL_L - 1][lane] = (uint16_t)lane;
                    }
                    wsync();
                    int sz = nc, cur = 1;
                    for (int l = CL_L - 2; l >= 0; l--) {
                        pm_level(lane, nc, sz, S.sf, S.cla[cur], S.cla[cur ^ 1], S.cllc[l]);
                        wsync();
                        sz = nc + (sz >> 1);
                        cur ^= 1;
                    }
                    if (lane == 0) {
                        int c = 2 * nc - 2;
                        for (int l = 0; l < CL_L; l++) {
                            const int xl = S.cllc[l][c];
                            S.clx[l] = (uint32_t)xl;
                            c = 2 * (c - xl);
                        }
                    }
                    wsync();
                    if (lane < nc) {
                        int len = 0;
                        for (int l = 0; l < CL_L; l++) len += lane < (int)S.clx[l] ? 1 : 0;
                        S.cl_len[S.cl_sorted[lane]] = (uint8_t)len;
                        atomicAdd(&S.clblc[len], 1u);
                    }
                }
            }
            __syncthreads();
            if (t < NCL) {
                S.cl_code[t] = canon_code(S.cl_len, S.clblc, t);
                if (S.cl_len[cl_order(t)]) atomicMax(&S.misc[3], (uint32_t)(t + 1));
            }
            __syncthreads();
            // the code-length symbols' bits (one symbol per thread) and their offsets
            uint32_t hpart = 0, hv = 0;
            if ((uint32_t)t < ns) {
                const int sym = S.cl_sym[t];
                const uint32_t eb = sym == 16 ? 2 : sym == 17 ? 3 : sym == 18 ? 7 : 0;
                hv = S.cl_code[sym] | (uint32_t)S.cl_extra[t] << S.cl_len[sym];
                hpart = S.cl_len[sym] + eb;
            }
            uint32_t hcl;
            const uint32_t hoff = block_excl_scan_1024(hpart, S.wsum, hcl);
            const uint32_t hclen = max(4u, S.misc[3]);
            const uint32_t hb = 3 + 5 + 5 + 4 + 3 * hclen + hcl;
            // 6. bit offsets of the sub-blocks
            uint32_t dbits = 0, wpre = 0;
            for (int v = 0; v < DF_W; v++) {
                const uint32_t x = S.wbits[v];
                if (v < w) wpre += x;
                dbits += x;
            }
            const uint32_t total_bits = hb + dbits + S.lit_len[256];
            nbytes = (total_bits + 7) >> 3;
            if (nbytes < (uint32_t)bl + 5) {
                // 7. the header and end of block, then every token, each ORed in at its
                //    scanned bit offset: BFINAL = 1, BTYPE = 2 (dynamic), HLIT, HDIST, HCLEN,
                //    the code-length code's lengths in transmission order, the code lengths
                if (t == 0) {
                    or_bits(S.u.out, 0, 1u | 2u << 1 | (uint32_t)(hlit - 257) << 3 | (uint32_t)(hdist - 1) << 8 |
                                            (hclen - 4) << 13, 17);
                    or_bits(S.u.out, hb + dbits, S.lpk[256] & 0xFFFFu, (int)(S.lpk[256] >> 16));
                }
                if ((uint32_t)t < hclen) or_bits(S.u.out, 17 + 3 * t, S.cl_len[cl_order(t)], 3);
                or_bits(S.u.out, 17 + 3 * hclen + hoff, hv, (int)hpart);
                // each lane takes 8 consecutive positions per 512: one scan per 512 positions
                uint32_t off = hb + wpre;
                for (int r0 = sb0; r0 < sb1; r0 += 8 * 64) {
                    const int P0 = r0 + 8 * lane;
                    uint32_t tk[9];  // tk[8]: the next position, the length of a match starting at 7
                    if (P0 < sb1) {
                        const uint4 a = *reinterpret_cast<const uint4 *>(tok + P0);
                        const uint32_t wv[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
                        for (int k = 0; k < 8; k++) tk[k] = (wv[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
                        tk[8] = (tk[7] & TOK_MATCH) ? tok[P0 + 8] : 0u;  // inside the match, so below sb1
                    } else {
#pragma unroll
                        for (int k = 0; k < 9; k++) tk[k] = 0;
                    }
                    // per token: c1 = literal or length code + extra bits (at most 20) | their
                    // count << 20 | the distance part's count << 25; c2 = distance code + extra
                    uint32_t c1[8], c2[8];
                    uint32_t bits = 0;
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const uint32_t x = P0 + k < sb1 ? tk[k] : 0u;
                        uint32_t v1 = 0, v2 = 0, b1 = 0, b2 = 0;
                        if (x & TOK_MATCH) {
                            const uint32_t mk = S.mpk[tk[k + 1]];
                            uint32_t ds, de, dx;
                            dist_code(x & 0x3FFF, ds, de, dx);
                            const uint32_t dk = S.dpk[ds];
                            v1 = mk & 0xFFFFFFu;
                            b1 = mk >> 24;
                            v2 = (dk & 0xFFFFu) | dx << (dk >> 16);
                            b2 = (dk >> 16) + de;
                        } else if (x & TOK_LIT) {
                            const uint32_t lk = S.lpk[x & 255];
                            v1 = lk & 0xFFFFu;
                            b1 = lk >> 16;
                        }
                        c1[k] = v1 | b1 << 20 | b2 << 25;
                        c2[k] = v2;
                        bits += b1 + b2;
                    }
                    const uint32_t incl = wave_incl_scan(bits);
                    if (bits) {
                        BitWriter bw(S.u.out, off + incl - bits);
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            const int b1 = (int)((c1[k] >> 20) & 31), b2 = (int)(c1[k] >> 25);
                            if (b1) bw.put(c1[k] & 0xFFFFFu, b1);
                            if (b2) bw.put(c2[k], b2);
                        }
                        bw.finish();
                    }
                    off += __builtin_amdgcn_readlane(incl, 63);
                }
                __syncthreads();
                const int nw = (int)((total_bits + 31) >> 5);
                for (int k = t; k < nw; k += DF_T) slot[k] = S.u.out[k];
            } else {
                nbytes = 0xFFFFFFFFu;
            }
        }
        // 8. sizes and CRC (the body is in the block's slot)
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
    size_t slots, body, crc, bytes, off, partial, tok, total;
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
    s.tok = take((size_t)g * DF_IN * 2);
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
                                                              body, crc, bytes, (uint16_t *)(s + L.tok));
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
