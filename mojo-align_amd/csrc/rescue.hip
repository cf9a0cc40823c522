// Mate rescue for paired batches, in place on the merged records.
//
// Science (bit-exact target): fm_mate_rescue_kernel of the reference's native
// BAM driver, fq2bam-meth/src/linear_gpu_fm.mojo:1839-2011. For a pair with
// exactly one mate mapped, the unmapped mate is scanned gapless at every start
// within +-window (clamped to 32..1024) of the mapped mate's position on that
// mate's contig copy, the strand opposite the mate's first, then the mate's
// own; a window's mismatches count until they pass budget + 4 (budget =
// clamp(max_diff, 2, len / 2)), and the first window of least count within
// budget wins. With none within budget, a one-base deletion is tried at the
// first window of least count, split points 0, 4, 8, ... The rescued mate
// gets the contig, the window start, flag 16 for the reverse strand, no clips,
// nm and MAPQ 60 / 40 / 20 (nm 0 / <= 2 / more).
//
// Execution (gfx950, wave64): a wave per candidate pair, waves striding over
// the pairs; the read and its reverse complement as 2-bit words (+ N flags)
// in LDS; lanes take 64 consecutive window starts per round and count a
// window's mismatches with XOR / popcount over whole 64-bit words of the 2-bit
// reference (32 bases per op); the reference's sequential first-minimum is a
// wave minimum over (count, scan index). The deletion: lane g / 4 scores split
// point g from two mismatch masks (the window and the window one base on).
#include "ma_index.hpp"
#include "ma_sort.hpp"

using namespace ma;

namespace {

constexpr int RS_WPB = 4;       // waves per block
constexpr int RS_MAXW = 8;      // read words (reads up to 256 bases)
constexpr uint64_t NONE = ~0ull;

struct RescueWave {
    uint8_t codes[256];            // converted read codes (> 3 = N)
    uint64_t w[2][RS_MAXW + 1];    // [0] read, [1] reverse complement; 2-bit words
    uint64_t n[2][RS_MAXW + 1];    // N flags (low bit of the base's group)
    uint64_t m[2][RS_MAXW + 1];    // deletion: mismatch masks at the window and one base on
};

__device__ __forceinline__ uint8_t rs_code(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

__device__ __forceinline__ uint64_t rs_uni64(uint64_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}

// Orders LDS traffic between the lanes of one wave: compiler barrier + wait for
// outstanding LDS operations.
__device__ __forceinline__ void rs_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t rs_wave_min(uint64_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t x = __shfl_xor(v, o, 64);
        v = x < v ? x : v;
    }
    return v;
}

// The 32 bases of the reference from global base b on (2-bit words, MSB-first;
// the table ends in zero pad words, so word + 1 is always in bounds).
__device__ __forceinline__ uint64_t ref_word(const uint64_t *__restrict__ a, int64_t b) {
    const int64_t wi = b >> 5;
    const int sh = (int)(b & 31) * 2;
    const uint64_t x = a[wi];
    return sh == 0 ? x : (x << sh) | (a[wi + 1] >> (64 - sh));
}

// One bit per mismatching base (low bit of its group), bases j < len of word k.
__device__ __forceinline__ uint64_t mismatch_bits(const IndexView &ix, int64_t b, const uint64_t *q,
                                                  const uint64_t *qn, int k, int len) {
    const uint64_t r = ref_word(ix.ref2, b + 32 * k);
    const uint64_t x = q[k] ^ r;
    uint64_t d = (x | (x >> 1)) & LOW2;
    d |= qn[k];
    if (ix.ref_has_n) d |= ref_word(ix.refN, b + 32 * k);
    const int rem = len - 32 * k;
    return rem >= 32 ? d : (d & first_groups_mask(rem));
}

__global__ void __launch_bounds__(RS_WPB * 64) mate_rescue_kernel(
    IndexView ix, const uint8_t *__restrict__ bases, int64_t stride, const uint32_t *__restrict__ lens,
    int64_t n_pairs, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2, int32_t window, int32_t max_diff,
    int32_t *__restrict__ cid, uint32_t *__restrict__ pos, uint32_t *__restrict__ mapq, int32_t *__restrict__ flag,
    uint32_t *__restrict__ sl, uint32_t *__restrict__ sr, uint32_t *__restrict__ nm, uint32_t *__restrict__ iop,
    uint32_t *__restrict__ iat, uint32_t *__restrict__ ilen) {
    __shared__ RescueWave waves[RS_WPB];
    RescueWave &S = waves[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int64_t l_pac = ix.n_bases;
    const int64_t n_waves = (int64_t)gridDim.x * RS_WPB;
    for (int64_t p = (int64_t)blockIdx.x * RS_WPB + (threadIdx.x >> 6); p < n_pairs; p += n_waves) {
        const int64_t i1 = p, i2 = p + n_pairs;
        const int32_t c1 = (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)cid[i1]);
        const int32_t c2 = (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)cid[i2]);
        if ((c1 >= 0) == (c2 >= 0)) continue;
        const int64_t src = c1 >= 0 ? i1 : i2, dst = c1 >= 0 ? i2 : i1;
        const int64_t crid = c1 >= 0 ? c1 : c2;
        if (crid >= ix.n_contigs) continue;
        const int qlen = (int)__builtin_amdgcn_readfirstlane(lens[dst]);
        if (qlen < 24 || qlen > max_len) continue;  // a length past its row is never read
        const int64_t coff = (int64_t)ix.contig_off[crid];
        const int64_t clen = (int64_t)ix.contig_off[crid + 1] - coff;
        const int64_t mate_pos = (int64_t)__builtin_amdgcn_readfirstlane(pos[src]);
        const bool mate_rev = ((uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)flag[src]) & 16u) != 0;
        const int64_t win = window < 32 ? 32 : (window > 1024 ? 1024 : window);
        const int64_t lo = mate_pos - win < 0 ? 0 : mate_pos - win;
        const int64_t hi = mate_pos + win > clen - 24 ? clen - 24 : mate_pos + win;
        if (hi <= lo) continue;
        int budget = max_diff < 2 ? 2 : max_diff;
        if (budget > qlen / 2) budget = qlen / 2;
        const int W = (qlen + 31) >> 5;

        // the mate's bases, converted, as codes in LDS
        const int mode = dst < n_pairs ? bs_r1 : bs_r2;
        const uint8_t *row = bases + dst * stride;
        for (int j = lane; j < qlen; j += 64) {
            uint8_t c;
            if (encoding == MA_ENC_ASCII) c = rs_code(row[j]);
            else c = (uint8_t)((row[j >> 2] >> (6 - 2 * (j & 3))) & 3);
            if (c == 1 && mode == MA_BS_C2T) c = 3;
            if (c == 2 && mode == MA_BS_G2A) c = 0;
            S.codes[j] = c;
        }
        rs_wave_sync();
        // lane k < W: read word k; lane 32 + k: reverse-complement word k
        if ((lane & 31) < W) {
            const int k = lane & 31, s = lane >> 5;
            uint64_t w = 0, nf = 0;
            for (int j = 0; j < 32; j++) {
                const int i = 32 * k + j;
                if (i >= qlen) break;
                uint8_t c = s == 0 ? S.codes[i] : S.codes[qlen - 1 - i];
                if (c > 3) nf |= 1ull << (62 - 2 * j);
                else w |= (uint64_t)(s == 0 ? c : 3 - c) << (62 - 2 * j);
            }
            S.w[s][k] = w;
            S.n[s][k] = nf;
        }
        rs_wave_sync();

        // gapless scan: two passes, 64 starts per round, first minimum by (count, order)
        const int64_t span = hi - lo + 1;
        const uint64_t cap = (uint64_t)budget + 5;
        uint64_t best = NONE, scan = NONE;
        for (int pass = 0; pass < 2; pass++) {
            const int s = (pass == 0) == mate_rev ? 0 : 1;  // 1 = the reverse strand
            for (int64_t r0 = 0; r0 < span; r0 += 64) {
                const int64_t local = lo + r0 + lane;
                const int64_t b = coff + local;
                uint64_t kb = NONE, ks = NONE;
                if (r0 + lane < span && b + qlen <= l_pac) {
                    uint64_t cnt = 0;
                    for (int k = 0; k < W; k++) cnt += (uint64_t)__popcll(mismatch_bits(ix, b, S.w[s], S.n[s], k, qlen));
                    const uint64_t order = (uint64_t)pass * (uint64_t)span + (uint64_t)(r0 + lane);
                    ks = ((cnt < cap ? cnt : cap) << 32) | order;
                    if (cnt <= (uint64_t)budget) kb = (cnt << 32) | order;
                }
                kb = rs_wave_min(kb);
                ks = rs_wave_min(ks);
                best = kb < best ? kb : best;
                scan = ks < scan ? ks : scan;
            }
        }
        best = rs_uni64(best);
        scan = rs_uni64(scan);
        int64_t out_pos = 0, out_flag = 0, out_nm = -1;
        if (best != NONE) {
            const uint64_t order = best & 0xFFFFFFFFull;
            const int pass = order >= (uint64_t)span ? 1 : 0;
            out_pos = lo + (int64_t)(order - (uint64_t)pass * (uint64_t)span);
            out_flag = ((pass == 0) == mate_rev) ? 0 : 16;
            out_nm = (int64_t)(best >> 32);
        } else if (scan != NONE) {
            // one-base deletion at the first least window: bases before the
            // split against the window, after it against the window one base on
            const uint64_t order = scan & 0xFFFFFFFFull;
            const int pass = order >= (uint64_t)span ? 1 : 0;
            const int s = (pass == 0) == mate_rev ? 0 : 1;
            const int64_t local = lo + (int64_t)(order - (uint64_t)pass * (uint64_t)span);
            const int64_t b = coff + local;
            if (b + qlen + 1 <= l_pac) {
                if (lane < 2 * W) {
                    const int k = lane % W, sh = lane / W;
                    S.m[sh][k] = mismatch_bits(ix, b + sh, S.w[s], S.n[s], k, qlen);
                }
                rs_wave_sync();
                uint64_t bd = NONE;
                for (int g0 = 0; g0 <= qlen; g0 += 256) {
                    const int g = g0 + 4 * lane;
                    uint64_t kd = NONE;
                    if (g <= qlen) {
                        uint64_t cnt = 1;
                        for (int k = 0; k < W; k++) {
                            const int first = 32 * k;  // bases [first, first + 32) of word k
                            const uint64_t before = g >= first + 32 ? ~0ull : first_groups_mask(g - first);
                            cnt += (uint64_t)__popcll(S.m[0][k] & before) + (uint64_t)__popcll(S.m[1][k] & ~before);
                        }
                        if (cnt <= (uint64_t)budget) kd = (cnt << 32) | (uint64_t)g;
                    }
                    kd = rs_wave_min(kd);
                    bd = kd < bd ? kd : bd;
                }
                bd = rs_uni64(bd);
                if (bd != NONE) {
                    out_pos = local;
                    out_flag = s ? 16 : 0;
                    out_nm = (int64_t)(bd >> 32);
                }
            }
        }
        if (out_nm >= 0 && lane == 0) {
            cid[dst] = (int32_t)crid;
            pos[dst] = (uint32_t)out_pos;
            mapq[dst] = out_nm == 0 ? 60u : (out_nm <= 2 ? 40u : 20u);
            flag[dst] = (int32_t)out_flag;
            sl[dst] = 0u;
            sr[dst] = 0u;
            nm[dst] = (uint32_t)out_nm;
            if (iop) iop[dst] = 0u;
            if (iat) iat[dst] = 0u;
            if (ilen) ilen[dst] = 0u;
        }
        rs_wave_sync();
    }
}

}  // namespace

extern "C" {

int ma_mate_rescue_dev(ma_index *ix, void *stream, const uint8_t *d_bases, int64_t stride, const uint32_t *d_lens,
                       int64_t n_pairs, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                       int32_t window, int32_t max_diff, ma_recs *d_merged) {
    MA_REQUIRE(ix && d_merged && n_pairs >= 0, "ma_mate_rescue_dev: bad arguments");
    MA_REQUIRE(d_merged->cid && d_merged->pos && d_merged->mapq && d_merged->flag && d_merged->sl && d_merged->sr &&
                   d_merged->nm,
               "ma_mate_rescue_dev: merged record arrays must be non-null");
    MA_REQUIRE(encoding == MA_ENC_ASCII || encoding == MA_ENC_2BIT, "ma_mate_rescue_dev: bad encoding");
    MA_REQUIRE(max_len >= 0 && max_len <= 32 * RS_MAXW, "ma_mate_rescue_dev: reads longer than 256 bases");
    MA_REQUIRE(encoding == MA_ENC_ASCII ? stride >= max_len : stride * 4 >= max_len,
               "ma_mate_rescue_dev: stride smaller than a read row");
    if (n_pairs == 0 || window <= 0) return MA_OK;  // a window of 0 turns rescue off
    const IndexView v = index_view(ix);
    int64_t blocks = (n_pairs + RS_WPB - 1) / RS_WPB;
    const int64_t cap = (int64_t)ix->n_cu * 8;
    if (blocks > cap) blocks = cap;
    mate_rescue_kernel<<<(unsigned)blocks, RS_WPB * 64, 0, (hipStream_t)stream>>>(
        v, d_bases, stride, d_lens, n_pairs, max_len, encoding, bs_r1, bs_r2, window, max_diff, d_merged->cid,
        d_merged->pos,
        d_merged->mapq, d_merged->flag, d_merged->sl, d_merged->sr, d_merged->nm, d_merged->iop, d_merged->iat,
        d_merged->ilen);
    MA_LAUNCHED();
    return MA_OK;
}

}  // extern "C"
