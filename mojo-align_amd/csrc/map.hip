// Seed -> vote -> extend -> rescue -> high-occ tally, FW and RC, merged — one
// 64-lane wave per read, the whole batch in one launch.
//
// Science (bit-exact target): vote_extend_kernel and the per-batch host driver
// of the reference's parity engine, fq2bam-meth/src/linear_gpu_locate.mojo
// (encode :419-451, locate :453-490, vote :903-973, KEEP :975-1004, caps
// :1006-1024, extend :1037-1243, rescue :1245-1532, high-occ tally :1283-1345,
// resolve :1534-1585, MAPQ :1587-1607, RC codes :574-600, RC need :2356-2362,
// merge :2101-2113, unmapped-middle rule of _hit_from_gpu :229-234).
//
// Execution model (gfx950, wave64):
//   * a wave owns one read; its state (read words, seed order, the 256-entry
//     vote table, KEEP picks, 48-entry exact tally) lives in LDS (~4.7 KB);
//   * reads and reference are 2-bit packed (32 bases per u64, MSB-first) with
//     parallel N-flag words, so one lane scores a whole read against a
//     reference window with ~2W XOR/popcount ops (W = read words);
//   * lanes run in parallel over seeds (encode + one 64-byte hash-bucket probe
//     each), over the postings of one seed (vote / rescue / exact tally) and
//     over (pick, shift) extend candidates; ballots and prefix counts keep the
//     reference's first-seen table order, and first-minimum reductions keep its
//     sequential best-hit choice (picks are in non-increasing vote order, so
//     the vote tie-break of the extend never fires);
//   * the rare vote-table overflow (more than 256 distinct loci) runs the
//     reference's first-min replacement one posting at a time.
// No host round trip between passes: FW pass 0 / pass 1, the need_rc test, RC
// passes and the FW/RC merge all run inside the kernel.
#include <string.h>

#include "ma_index.hpp"
#include "ma_sort.hpp"

using namespace ma;

namespace {

constexpr int MAXW = 8;  // read words: reads up to 256 bases
constexpr int MAX_READ = MAXW * 32;
constexpr int TOP = 256;  // vote table (TOP, linear_gpu_locate.mojo:812)
constexpr int KEEP = 16;  // extend picks (KEEP, :813)
constexpr int HX = 48;    // exact tally (HX, :1249)
constexpr int MAXS = 64;  // seeds per read (one per lane)
constexpr int WPB = 4;    // waves per block
constexpr int BLOCK = WPB * 64;
constexpr int MAX_HIGH_TRIED = 96;  // :1277-1281

struct WaveState {
    uint64_t fw2[MAXW + 2];  // [0] = 0 pad, read words at [1..W], zero tail
    uint64_t fwN[MAXW + 2];
    uint64_t rd2[MAXW + 1];  // read of the current pass (FW or RC), zero tail
    uint64_t rdN[MAXW + 1];
    uint64_t ord[MAXS];  // occ << 46 | slot << 40 | start, ascending
    uint64_t cand_key[TOP];
    int32_t cand_n[TOP];
    int32_t pick_i[KEEP];
    int32_t pick_n[KEEP];
    uint64_t hx_key[HX];
    int32_t hx_n[HX];
    int32_t hx_min[HX];
};

struct Knobs {
    int32_t k, stride, max_occ, vote_occ, vote_fast, max_diff, max_soft, do_rc;
};

struct Rec {
    int32_t cid;
    uint32_t pos, mapq;
    int32_t flag;
    uint32_t sl, sr, nm;
};

struct RecOut {
    int32_t *cid;
    uint32_t *pos, *mapq;
    int32_t *flag;
    uint32_t *sl, *sr, *nm, *iop, *iat, *ilen;
};

struct ReadsIn {
    const uint8_t *bases;
    int64_t stride;
    const uint32_t *lens;
    int64_t n;
    int32_t max_len;
    int32_t encoding;
    int32_t bs_r1, bs_r2;
    int64_t n_r1;
};

__device__ inline Rec unmapped_rec() { return Rec{-1, 0u, 0u, 4, 0u, 0u, 0u}; }

__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }

// Orders LDS traffic between lanes of one wave (lockstep lanes, in-order LDS
// queue per wave): compiler barrier + wait for outstanding LDS operations.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint64_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// Multiset top-2 (b1 >= b2) across the wave.
__device__ inline void wave_top2(int &b1, int &b2) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        int c1 = __shfl_xor(b1, o, 64), c2 = __shfl_xor(b2, o, 64);
        int hi = b1 > c1 ? b1 : c1;
        int lo = b1 > c1 ? c1 : b1;
        int s = b2 > c2 ? b2 : c2;
        b1 = hi;
        b2 = lo > s ? lo : s;
    }
}

__device__ inline void top2_push(int v, int &b1, int &b2) {
    if (v > b1) {
        b2 = b1;
        b1 = v;
    } else if (v > b2) {
        b2 = v;
    }
}

// 32 bases at base offset p of a zero-padded LDS word array.
__device__ inline uint64_t lds_window(const uint64_t *w, int p) {
    int wi = p >> 5, sh = (p & 31) * 2;
    uint64_t a = w[wi];
    if (sh == 0) return a;
    return (a << sh) | (w[wi + 1] >> (64 - sh));
}

__device__ inline uint64_t pick_word(const uint64_t (&m)[MAXW], int i) {
    uint64_t v = 0;
#pragma unroll
    for (int t = 0; t < MAXW; t++) v = (t == i) ? m[t] : v;
    return v;
}

// Both bits of the groups j < n.
__device__ inline uint64_t groups_mask2(int n) {
    uint64_t f = first_groups_mask(n);
    return f | (f << 1);
}

// Mismatch flags (low bit of each 2-bit group) of the current read against
// the reference window starting at global base g; bases >= L read as matches.
// A base mismatches when either side is N or the codes differ.
__device__ inline int window_mm(const IndexView &ix, int64_t g, const WaveState &S, int W, int L,
                                uint64_t (&mm)[MAXW]) {
    int tot = 0;
#pragma unroll
    for (int t = 0; t < MAXW; t++) {
        uint64_t m = 0;
        if (t < W) {
            uint64_t r = window32(ix.ref2, g + 32 * t);
            uint64_t x = r ^ S.rd2[t];
            m = (x | (x >> 1)) & LOW2;
            if (ix.ref_has_n) m |= window32(ix.refN, g + 32 * t);
            m |= S.rdN[t];
            m &= first_groups_mask(L - 32 * t);
        }
        mm[t] = m;
        tot += __popcll(m);
    }
    return tot;
}

__device__ inline bool mm_bit(const uint64_t (&mm)[MAXW], int j) {
    return (pick_word(mm, j >> 5) >> (62 - 2 * (j & 31))) & 1ull;
}

// Index of the first matching base (L if none).
__device__ inline int first_match(const uint64_t (&mm)[MAXW], int W, int L) {
    int res = L;
    bool found = false;
#pragma unroll
    for (int t = 0; t < MAXW; t++) {
        if (!found && t < W) {
            uint64_t inv = ~mm[t] & first_groups_mask(L - 32 * t);
            if (inv) {
                res = 32 * t + (__clzll(inv) - 1) / 2;
                found = true;
            }
        }
    }
    return res;
}

// Number of consecutive mismatching bases ending at base L-1.
__device__ inline int trailing_run(const uint64_t (&mm)[MAXW], int L) {
    int run = 0;
    int t = (L - 1) >> 5;
    int gl = (L - 1) & 31;
    while (t >= 0) {
        uint64_t v = pick_word(mm, t) >> (62 - 2 * gl);
        uint64_t span = (gl == 31) ? ~0ull : ((1ull << (2 * gl + 2)) - 1);
        uint64_t inv = ~v & LOW2 & span;
        if (inv) return run + (__ffsll((long long)inv) - 1) / 2;
        run += gl + 1;
        t--;
        gl = 31;
    }
    return run;
}

// Mismatches among bases [a, b).
__device__ inline int count_range(const uint64_t (&mm)[MAXW], int a, int b) {
    int tot = 0;
#pragma unroll
    for (int t = 0; t < MAXW; t++) {
        uint64_t m = mm[t] & first_groups_mask(b - 32 * t) & ~first_groups_mask(a - 32 * t);
        tot += __popcll(m);
    }
    return tot;
}

struct Caps {
    int budget, clip_cap, shift_cap;
};

__device__ inline Caps make_caps(const Knobs &kn, int L) {
    Caps c;
    int budget = kn.max_diff;
    if (budget < 2) budget = 2;
    if (budget > L / 2) budget = L / 2;
    int clip = kn.max_soft;
    if (clip > L / 4) clip = L / 4;
    c.budget = budget;
    c.clip_cap = clip;
    c.shift_cap = 2;  // clamp(indel_cap = 0, 2, 6)
    return c;
}

// Soft-clip both ends (sl, sr <= clip_cap, sr stops before sl) and recount
// with the reference's early abort (the count saturates at budget + 1).
__device__ inline int softclip(const uint64_t (&mm)[MAXW], int W, int L, const Caps &cp, int &sl,
                               int &sr) {
    int fm = first_match(mm, W, L);
    sl = fm < cp.clip_cap ? fm : cp.clip_cap;
    int tr = trailing_run(mm, L);
    int lim = L - 1 - sl;
    if (lim < 0) lim = 0;
    sr = tr < cp.clip_cap ? tr : cp.clip_cap;
    if (sr > lim) sr = lim;
    int nm = count_range(mm, sl, L - sr);
    return nm > cp.budget ? cp.budget + 1 : nm;
}

__device__ inline uint64_t load_posting(const IndexView &ix, int64_t pi) {
    // (cid, pos) u32 pair: cid in the low word
    return *(const uint64_t *)(ix.postings + 2 * pi);
}

// ------------------------------------------------------------------ seeds
// Lane s encodes slot s (q_off = s * stride) and probes the hash bucket;
// valid seeds (0 < occ <= max_occ) are ranked by (occ, slot) into S.ord.
__device__ int seed_and_order(const IndexView &ix, const Knobs &kn, WaveState &S, int L) {
    int lane = lane_id();
    int k = kn.k;
    int n_slots = L >= k ? (L - k) / kn.stride + 1 : 0;
    uint64_t e = ~0ull;
    bool valid = false;
    if (lane < n_slots) {
        int q = lane * kn.stride;
        uint64_t nw = lds_window(S.rdN, q) >> (64 - 2 * k);
        if (nw == 0) {
            uint64_t key = lds_window(S.rd2, q) >> (64 - 2 * k);
            uint64_t b = bucket_of(key, ix.n_buckets);
            bool done = false;
            for (int probe = 0; probe < ix.max_probe && !done; probe++) {
                const HashSlot *bk = ix.table + b * BUCKET_SLOTS;
#pragma unroll
                for (int s = 0; s < BUCKET_SLOTS; s++) {
                    if (done) continue;
                    uint64_t kk = bk[s].key;
                    if (kk == key) {
                        uint64_t val = bk[s].val;
                        uint64_t occ = val >> 40;
                        if (occ >= 1 && occ <= (uint64_t)kn.max_occ) {
                            e = (occ << 46) | ((uint64_t)lane << 40) | (val & START_MASK);
                            valid = true;
                        }
                        done = true;
                    } else if (kk == EMPTY_KEY) {
                        done = true;
                    }
                }
                b = (b + 1 == ix.n_buckets) ? 0 : b + 1;
            }
        }
    }
    uint64_t vm = __ballot(valid);
    int n_ord = __popcll(vm);
    int rank = 0;
    uint64_t m = vm;
    while (m) {
        int j = __ffsll((long long)m) - 1;
        m &= m - 1;
        uint64_t ej = __shfl(e, j, 64);
        rank += (ej < e) ? 1 : 0;
    }
    if (valid) S.ord[rank] = e;
    wave_sync();
    return n_ord;
}

__device__ inline Rec finish_rec(int64_t cid, int64_t pos, int votes, int cost, int sl, int sr,
                                 int32_t strand_flag) {
    Rec r;
    uint32_t mq = 20;
    if (votes >= 3) mq = 40;
    if (votes >= 5) mq = 60;
    if (cost > 2 && mq > 20) mq = 20;
    r.cid = (int32_t)cid;
    r.pos = (uint32_t)pos;
    r.mapq = mq;
    r.flag = strand_flag;
    r.sl = (uint32_t)sl;
    r.sr = (uint32_t)sr;
    r.nm = (uint32_t)cost;
    return r;
}

// ----------------------------------------------------------------- pass 0
// Vote over seeds with occ <= vote_cap, KEEP, extend at shifts 0,-1,+1,-2,+2,
// then (if nothing accepted) the low-occ rescue.
__device__ Rec pass0(const IndexView &ix, const Knobs &kn, WaveState &S, int n_ord, int W, int L,
                     int vote_cap_in, int32_t strand_flag) {
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt(lane);
    const int vote_cap = vote_cap_in < 1 ? 256 : vote_cap_in;
    int n_cand = 0;

    for (int oi = 0; oi < n_ord; oi++) {
        uint64_t e = S.ord[oi];
        int occ = (int)(e >> 46);
        if (occ > vote_cap) break;  // ord is ascending: no later seed votes
        int q_off = (int)((e >> 40) & 63) * kn.stride;
        int64_t start = (int64_t)(e & START_MASK);
        for (int c0 = 0; c0 < occ; c0 += 64) {
            bool valid = false;
            uint64_t vk = 0;
            if (c0 + lane < occ) {
                uint64_t p = load_posting(ix, start + c0 + lane);
                uint32_t cid = (uint32_t)p, pos = (uint32_t)(p >> 32);
                if ((int64_t)cid < ix.n_contigs && (int64_t)pos >= q_off) {
                    valid = true;
                    vk = ((uint64_t)cid << 32) | (uint64_t)(pos - (uint32_t)q_off);
                }
            }
            uint64_t vm = __ballot(valid);
            if (!vm) continue;
            int found = -1;
            for (int c = 0; c < n_cand; c++) {
                uint64_t ck = S.cand_key[c];
                if (valid && ck == vk) found = c;
            }
            bool isnew = valid && found < 0;
            uint64_t nwm = __ballot(isnew);
            int n_new = __popcll(nwm);
            if (n_cand + n_new <= TOP) {
                // keys of one seed are distinct: counts and appends never collide
                if (valid && found >= 0) S.cand_n[found] += 1;
                if (isnew) {
                    int r = __popcll(nwm & lt);
                    S.cand_key[n_cand + r] = vk;
                    S.cand_n[n_cand + r] = 1;
                }
                n_cand += n_new;
                wave_sync();
            } else {
                // table would overflow: postings one at a time, in order
                uint64_t m = vm;
                while (m) {
                    int j = __ffsll((long long)m) - 1;
                    m &= m - 1;
                    uint64_t key = __shfl(vk, j, 64);
                    int f = -1;
                    for (int base = 0; base < n_cand; base += 64) {
                        int c = base + lane;
                        bool hit = c < n_cand && S.cand_key[c] == key;
                        uint64_t hm = __ballot(hit);
                        if (hm) {
                            f = base + __ffsll((long long)hm) - 1;
                            break;
                        }
                    }
                    if (f >= 0) {
                        if (lane == 0) S.cand_n[f] += 1;
                    } else if (n_cand < TOP) {
                        if (lane == 0) {
                            S.cand_key[n_cand] = key;
                            S.cand_n[n_cand] = 1;
                        }
                        n_cand++;
                    } else {
                        uint64_t best = ~0ull;
                        for (int base = 0; base < TOP; base += 64) {
                            int c = base + lane;
                            uint64_t v = ((uint64_t)(uint32_t)S.cand_n[c] << 32) | (uint64_t)c;
                            best = v < best ? v : best;
                        }
                        best = wave_min_u64(best);
                        int min_n = (int)(best >> 32);
                        int min_i = (int)(best & 0xFFFFFFFFull);
                        if (min_n <= 1 && lane == 0) {
                            S.cand_key[min_i] = key;
                            S.cand_n[min_i] = 1;
                        }
                    }
                    wave_sync();
                }
            }
        }
        if (n_cand > 0 && oi + 1 < n_ord) {
            int b1 = 0, b2 = 0;
            for (int c = lane; c < n_cand; c += 64) top2_push(S.cand_n[c], b1, b2);
            wave_top2(b1, b2);
            if (b1 >= 4 && b1 >= b2 + 2) break;
        }
    }

    // KEEP: stable top-16 by votes
    int n_pick = n_cand < KEEP ? n_cand : KEEP;
    for (int base = 0; base < n_cand; base += 64) {
        int c = base + lane;
        if (c < n_cand) {
            int v = S.cand_n[c];
            int rank = 0;
            for (int c2 = 0; c2 < n_cand; c2++) {
                int v2 = S.cand_n[c2];
                rank += (v2 > v || (v2 == v && c2 < c)) ? 1 : 0;
            }
            if (rank < KEEP) {
                S.pick_i[rank] = c;
                S.pick_n[rank] = v;
            }
        }
    }
    wave_sync();

    const Caps cp = make_caps(kn, L);
    const int NA = 2 * cp.shift_cap + 1;
    const int NE = n_pick * NA;
    uint64_t best = ~0ull;  // cost << 48 | order << 16 | sl << 8 | sr
    int e_begin = 0, width = 1;
    while (e_begin < NE) {
        int ev = e_begin + lane;
        uint64_t cand = ~0ull;
        if (lane < width && ev < NE) {
            int pk = ev / NA, ai = ev - pk * NA;
            int mag = (ai + 1) >> 1;
            int adj = ai == 0 ? 0 : ((ai & 1) ? -mag : mag);
            uint64_t key = S.cand_key[S.pick_i[pk]];
            int64_t bcid = (int64_t)(key >> 32);
            int64_t bstart = (int64_t)(key & 0xFFFFFFFFull) + adj;
            if (bcid < ix.n_contigs) {
                int64_t off0 = (int64_t)ix.contig_off[bcid];
                int64_t clen = (int64_t)ix.contig_off[bcid + 1] - off0;
                if (bstart >= 0 && bstart + L <= clen) {
                    uint64_t mm[MAXW];
                    int tot = window_mm(ix, off0 + bstart, S, W, L, mm);
                    int probe_nm = 0;
#pragma unroll
                    for (int p = 1; p <= 6; p++) {
                        int jp = (p * (L - 1)) / 7;
                        if (jp < cp.clip_cap) jp = cp.clip_cap;
                        if (jp > L - 1 - cp.clip_cap) jp = L - 1 - cp.clip_cap;
                        if (jp >= 0 && jp < L && mm_bit(mm, jp)) probe_nm++;
                    }
                    if (probe_nm <= cp.budget) {
                        int nm = tot > cp.budget ? cp.budget + 1 : tot;
                        int sl = 0, sr = 0;
                        if (nm > cp.budget) nm = softclip(mm, W, L, cp, sl, sr);
                        if (nm <= cp.budget && L - sl - sr >= 24)
                            cand = ((uint64_t)nm << 48) | ((uint64_t)ev << 16) | ((uint64_t)sl << 8) |
                                   (uint64_t)sr;
                    }
                }
            }
        }
        uint64_t mn = wave_min_u64(cand);
        if (mn < best) best = mn;
        if (best != ~0ull && (best >> 48) == 0) break;
        e_begin += width;
        width = 64;
    }
    if (best != ~0ull) {
        int cost = (int)(best >> 48);
        int ev = (int)((best >> 16) & 0xFFFF);
        int sl = (int)((best >> 8) & 0xFF), sr = (int)(best & 0xFF);
        int pk = ev / NA, ai = ev - pk * NA;
        int mag = (ai + 1) >> 1;
        int adj = ai == 0 ? 0 : ((ai & 1) ? -mag : mag);
        uint64_t key = S.cand_key[S.pick_i[pk]];
        int64_t pos = (int64_t)(key & 0xFFFFFFFFull) + adj + sl;
        return finish_rec((int64_t)(key >> 32), pos, S.pick_n[pk], cost, sl, sr, strand_flag);
    }

    // low-occ rescue: every posting of every seed with occ <= vote_cap, first
    // minimum cost wins, an exact hit ends the scan
    int best_cost = 1 << 30;
    int64_t b_cid = -1, b_pos = 0;
    int b_sl = 0, b_sr = 0, b_votes = 0;
    for (int oi = 0; oi < n_ord; oi++) {
        uint64_t e = S.ord[oi];
        int occ = (int)(e >> 46);
        if (occ > vote_cap) break;
        int q_off = (int)((e >> 40) & 63) * kn.stride;
        int64_t start = (int64_t)(e & START_MASK);
        bool stop = false;
        for (int c0 = 0; c0 < occ; c0 += 64) {
            uint64_t cand = ~0ull;
            int64_t my_cid = 0, my_pos = 0;
            if (c0 + lane < occ) {
                uint64_t p = load_posting(ix, start + c0 + lane);
                uint32_t cid = (uint32_t)p, pos = (uint32_t)(p >> 32);
                if ((int64_t)cid < ix.n_contigs && (int64_t)pos >= q_off) {
                    int64_t bstart = (int64_t)pos - q_off;
                    int64_t off0 = (int64_t)ix.contig_off[cid];
                    int64_t clen = (int64_t)ix.contig_off[cid + 1] - off0;
                    if (bstart + L <= clen) {
                        uint64_t mm[MAXW];
                        int tot = window_mm(ix, off0 + bstart, S, W, L, mm);
                        int cost = tot > cp.budget ? cp.budget + 1 : tot;
                        int sl = 0, sr = 0;
                        int64_t pos_out = bstart;
                        if (cost > cp.budget && cp.clip_cap > 0) {
                            cost = softclip(mm, W, L, cp, sl, sr);
                            pos_out = bstart + sl;
                            if (L - sl - sr < 32) cost = 9999;
                        }
                        if (cost <= cp.budget) {
                            cand = ((uint64_t)cost << 48) | ((uint64_t)lane << 16) |
                                   ((uint64_t)sl << 8) | (uint64_t)sr;
                            my_cid = cid;
                            my_pos = pos_out;
                        }
                    }
                }
            }
            uint64_t mn = wave_min_u64(cand);
            if (mn != ~0ull) {
                int cost = (int)(mn >> 48);
                int wl = (int)((mn >> 16) & 0xFFFF);
                int64_t wc = __shfl(my_cid, wl, 64);
                int64_t wp = __shfl(my_pos, wl, 64);
                if (cost < best_cost) {
                    best_cost = cost;
                    b_cid = wc;
                    b_pos = wp;
                    b_sl = (int)((mn >> 8) & 0xFF);
                    b_sr = (int)(mn & 0xFF);
                    b_votes = occ <= 64 ? 5 : 4;
                }
                if (best_cost == 0) {
                    stop = true;
                    break;
                }
            }
        }
        if (stop) break;
    }
    if (b_cid >= 0) return finish_rec(b_cid, b_pos, b_votes, best_cost, b_sl, b_sr, strand_flag);
    return unmapped_rec();
}

// ----------------------------------------------------------------- pass 1
// High-occ exact tally for a read left unmapped by pass 0.
__device__ Rec pass1(const IndexView &ix, const Knobs &kn, WaveState &S, int n_ord, int W, int L,
                     int vote_cap_in, int32_t strand_flag) {
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt(lane);
    const int vote_cap = vote_cap_in < 1 ? 256 : vote_cap_in;
    int n_hx = 0, tried = 0;
    for (int oi = 0; oi < n_ord; oi++) {
        uint64_t e = S.ord[oi];
        int occ = (int)(e >> 46);
        if (occ <= vote_cap) continue;
        tried++;
        if (tried > MAX_HIGH_TRIED) break;
        int q_off = (int)((e >> 40) & 63) * kn.stride;
        int64_t start = (int64_t)(e & START_MASK);
        for (int c0 = 0; c0 < occ; c0 += 64) {
            bool exact = false;
            uint64_t vk = 0;
            if (c0 + lane < occ) {
                uint64_t p = load_posting(ix, start + c0 + lane);
                uint32_t cid = (uint32_t)p, pos = (uint32_t)(p >> 32);
                if ((int64_t)cid < ix.n_contigs && (int64_t)pos >= q_off) {
                    int64_t bs = (int64_t)pos - q_off;
                    int64_t o0 = (int64_t)ix.contig_off[cid];
                    int64_t clen = (int64_t)ix.contig_off[cid + 1] - o0;
                    if (bs + L <= clen) {
                        uint64_t mm[MAXW];
                        if (window_mm(ix, o0 + bs, S, W, L, mm) == 0) {
                            exact = true;
                            vk = ((uint64_t)cid << 32) | (uint64_t)bs;
                        }
                    }
                }
            }
            uint64_t xm = __ballot(exact);
            if (!xm) continue;
            int found = -1;
            for (int h = 0; h < n_hx; h++) {
                uint64_t hk = S.hx_key[h];
                if (exact && hk == vk) found = h;
            }
            bool isnew = exact && found < 0;
            uint64_t nwm = __ballot(isnew);
            if (exact && found >= 0) {
                S.hx_n[found] += 1;
                if (occ < S.hx_min[found]) S.hx_min[found] = occ;
            }
            if (isnew) {
                int r = n_hx + __popcll(nwm & lt);
                if (r < HX) {
                    S.hx_key[r] = vk;
                    S.hx_n[r] = 1;
                    S.hx_min[r] = occ;
                }
            }
            n_hx += __popcll(nwm);
            if (n_hx > HX) n_hx = HX;
            wave_sync();
        }
        if (n_hx > 0) {
            int b1 = 0, b2 = 0;
            for (int h = lane; h < n_hx; h += 64) top2_push(S.hx_n[h], b1, b2);
            wave_top2(b1, b2);
            if (b1 >= 3 && b1 > b2) break;
        }
    }
    if (n_hx == 0) return unmapped_rec();
    int best_hi = 0, second_hi = 0, best_i = -1, best_mocc = 0x7FFFFFFF;
    for (int h = 0; h < n_hx; h++) {
        int vn = S.hx_n[h], mo = S.hx_min[h];
        if (vn > best_hi) {
            second_hi = best_hi;
            best_hi = vn;
            best_mocc = mo;
            best_i = h;
        } else if (vn == best_hi && mo < best_mocc) {
            best_mocc = mo;
            best_i = h;
        } else if (vn > second_hi) {
            second_hi = vn;
        }
    }
    if (best_i >= 0 && (best_hi >= 2 || (best_hi >= 1 && second_hi == 0))) {
        uint64_t bk = S.hx_key[best_i];
        int votes = best_hi > second_hi ? (best_hi > 5 ? 5 : best_hi) : 1;
        return finish_rec((int64_t)(bk >> 32), (int64_t)(bk & 0xFFFFFFFFull), votes, 0, 0, 0,
                          strand_flag);
    }
    return unmapped_rec();
}

__device__ Rec map_strand(const IndexView &ix, const Knobs &kn, WaveState &S, int W, int L,
                          int32_t strand_flag) {
    int n_ord = seed_and_order(ix, kn, S, L);
    Rec r = pass0(ix, kn, S, n_ord, W, L, kn.vote_fast, strand_flag);
    if (r.cid < 0 && kn.vote_fast < kn.vote_occ)
        r = pass0(ix, kn, S, n_ord, W, L, kn.vote_occ, strand_flag);
    if (r.cid < 0) r = pass1(ix, kn, S, n_ord, W, L, kn.vote_occ, strand_flag);
    return r;
}

__device__ inline uint8_t ascii_code(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 255;
    }
}

// Lane t < W builds read word t (+ N word), applying the in-silico conversion.
__device__ void load_read(const ReadsIn &in, int64_t r, int L, int W, WaveState &S) {
    int lane = lane_id();
    int mode = r < in.n_r1 ? in.bs_r1 : in.bs_r2;
    const uint8_t *src = in.bases + r * in.stride;
    uint64_t w = 0, nf = 0;
    if (lane < W) {
        if (in.encoding == MA_ENC_ASCII) {
            for (int j = 0; j < 32; j++) {
                int i = 32 * lane + j;
                if (i >= L) break;
                uint8_t c = ascii_code(src[i]);
                if (c == 1 && mode == MA_BS_C2T) c = 3;
                if (c == 2 && mode == MA_BS_G2A) c = 0;
                if (c > 3) nf |= 1ull << (62 - 2 * j);
                else w |= (uint64_t)c << (62 - 2 * j);
            }
        } else {
            int nbytes = (L + 3) >> 2;
            for (int j = 0; j < 8; j++) {
                int bi = 8 * lane + j;
                if (bi < nbytes) w |= (uint64_t)src[bi] << (56 - 8 * j);
            }
            if (mode == MA_BS_C2T) w |= (w & LOW2) << 1;
            if (mode == MA_BS_G2A) w &= ~((~w & LOW2) << 1);
            w &= groups_mask2(L - 32 * lane);
        }
    }
    if (lane < MAXW + 2) {
        S.fw2[lane] = 0;
        S.fwN[lane] = 0;
    }
    wave_sync();
    if (lane < W) {
        S.fw2[lane + 1] = w;
        S.fwN[lane + 1] = nf;
    }
    if (lane < MAXW + 1) {
        S.rd2[lane] = lane < W ? w : 0;
        S.rdN[lane] = lane < W ? nf : 0;
    }
    wave_sync();
}

// Reverse complement of the FW read into rd2/rdN (N stays N, codes c -> 3-c).
__device__ void make_rc(WaveState &S, int L, int W) {
    int lane = lane_id();
    uint64_t w = 0, nf = 0;
    if (lane < W) {
        int p = L - 32 * (lane + 1) + 32;  // position in the padded array
        uint64_t mask2 = groups_mask2(L - 32 * lane);
        w = ~reverse_groups(lds_window(S.fw2, p)) & mask2;
        nf = reverse_groups(lds_window(S.fwN, p)) & first_groups_mask(L - 32 * lane);
        // an N base carries code 0 in fw2; its complement bits are irrelevant
        // (the N flag makes it a mismatch) but keep them zero like FW
        w &= ~(nf | (nf << 1));
    }
    wave_sync();
    if (lane < MAXW + 1) {
        S.rd2[lane] = lane < W ? w : 0;
        S.rdN[lane] = lane < W ? nf : 0;
    }
    wave_sync();
}

__device__ inline void store_rec(const RecOut &o, int64_t r, const Rec &x) {
    if (!o.cid) return;
    o.cid[r] = x.cid;
    o.pos[r] = x.pos;
    o.mapq[r] = x.mapq;
    o.flag[r] = x.flag;
    o.sl[r] = x.sl;
    o.sr[r] = x.sr;
    o.nm[r] = x.nm;
    if (o.iop) o.iop[r] = 0;
    if (o.iat) o.iat[r] = 0;
    if (o.ilen) o.ilen[r] = 0;
}

__global__ void __launch_bounds__(BLOCK) map_reads_kernel(IndexView ix, Knobs kn, ReadsIn in,
                                                           RecOut merged, RecOut fw_out,
                                                           RecOut rc_out,
                                                           unsigned int *__restrict__ counter) {
    __shared__ WaveState states[WPB];
    WaveState &S = states[threadIdx.x >> 6];
    const int lane = lane_id();
    while (true) {
        unsigned int rr = 0;
        if (lane == 0) rr = atomicAdd(counter, 1u);
        int64_t r = (int64_t)__shfl((int)rr, 0, 64);
        r = (int64_t)(uint32_t)r;
        if (r >= in.n) break;
        int L = (int)in.lens[r];
        if (L > in.max_len) L = 0;  // malformed row: reported unmapped, never read past its row
        Rec fw = unmapped_rec(), rc = unmapped_rec();
        if (L > 0 && L <= MAX_READ) {
            int W = (L + 31) >> 5;
            load_read(in, r, L, W, S);
            fw = map_strand(ix, kn, S, W, L, 0);
            bool need_rc = kn.do_rc && (fw.cid < 0 || fw.nm > 0);
            if (need_rc) {
                make_rc(S, L, W);
                rc = map_strand(ix, kn, S, W, L, 16);
            }
        }
        bool use_rc = false;
        if (fw.cid < 0 && rc.cid >= 0) use_rc = true;
        else if (fw.cid >= 0 && rc.cid >= 0 &&
                 (rc.nm < fw.nm || (rc.nm == fw.nm && rc.mapq > fw.mapq)))
            use_rc = true;
        Rec m = use_rc ? rc : fw;
        if (m.cid < 0 || L - (int)m.sl - (int)m.sr < 1) m = unmapped_rec();
        if (lane == 0) {
            store_rec(merged, r, m);
            store_rec(fw_out, r, fw);
            store_rec(rc_out, r, rc);
        }
    }
}

__global__ void zero_u32_kernel(unsigned int *p) { *p = 0u; }

// gpu-common seed kernel: per position, Wang hash of the MSB-first k-mer
// (0 when the window passes the row end or holds a non-ACGT base).
// Follows kmer_hash_kernel, gpu-common/src/gpu_kernels.mojo:139-170.
__global__ void seed_hash_kernel(const uint8_t *__restrict__ bases, int64_t n_bases, int64_t row,
                                 int k, uint64_t *__restrict__ out) {
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n_bases;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t pos = idx % row;
        uint64_t h = 0;
        if (pos + k <= row) {
            const uint8_t *p = bases + (idx / row) * row + pos;
            uint64_t key = 0;
            bool ok = true;
            for (int j = 0; j < k; j++) {
                uint8_t c = ascii_code(p[j]);
                if (c > 3) {
                    ok = false;
                    break;
                }
                key = (key << 2) | c;
            }
            if (ok) h = wang_hash(key);
        }
        out[idx] = h;
    }
}

RecOut rec_out(const ma_recs *r) {
    RecOut o{};
    if (!r) return o;
    o.cid = r->cid;
    o.pos = r->pos;
    o.mapq = r->mapq;
    o.flag = r->flag;
    o.sl = r->sl;
    o.sr = r->sr;
    o.nm = r->nm;
    o.iop = r->iop;
    o.iat = r->iat;
    o.ilen = r->ilen;
    if (!o.cid || !o.pos || !o.mapq || !o.flag || !o.sl || !o.sr || !o.nm) o.cid = nullptr;
    return o;
}

int check_params(const ma_index *ix, const ma_params *p, int32_t max_len) {
    MA_REQUIRE(p, "map: null params");
    MA_REQUIRE(p->k == ix->k, "map: params.k differs from the index k");
    MA_REQUIRE(p->stride >= 1, "map: stride must be >= 1");
    MA_REQUIRE(p->max_indel == 0,
               "map: max_indel must be 0 (gapless extension; indel branches are not built)");
    MA_REQUIRE(p->max_occ >= 1 && p->max_occ < (1 << 17), "map: max_occ must be in 1..131071");
    MA_REQUIRE(max_len >= 0 && max_len <= MAX_READ, "map: reads longer than 256 bases are not supported");
    int32_t ml = max_len < p->k ? p->k : max_len;
    MA_REQUIRE((ml - p->k) / p->stride + 1 <= MAXS, "map: more than 64 seed slots per read");
    return MA_OK;
}

}  // namespace

extern "C" {

int64_t ma_workspace_bytes(int64_t max_reads, int32_t max_len) {
    (void)max_reads;
    (void)max_len;
    return 256;
}

int ma_workspace_create(ma_index *ix, int64_t max_reads, int32_t max_len, ma_workspace **out) {
    MA_REQUIRE(ix && out && max_reads >= 0 && max_len >= 0, "ma_workspace_create: bad arguments");
    MA_HIP(hipSetDevice(ix->device));
    ma_workspace *ws = new ma_workspace();
    ws->ix = ix;
    ws->max_reads = max_reads;
    ws->max_len = max_len;
    int rc = ws->counter.alloc(256);
    if (rc != MA_OK) {
        delete ws;
        return rc;
    }
    *out = ws;
    return MA_OK;
}

int ma_workspace_destroy(ma_workspace *ws) {
    delete ws;
    return MA_OK;
}

int ma_map_batch_dev(ma_index *ix, ma_workspace *ws, void *stream, const uint8_t *d_bases,
                     int64_t stride, const uint32_t *d_lens, int64_t n, int32_t max_len,
                     int32_t encoding, int32_t bs_r1, int32_t bs_r2, int64_t n_r1,
                     const ma_params *p, ma_recs *d_merged, ma_recs *d_fw, ma_recs *d_rc) {
    MA_REQUIRE(ix && ws && d_merged, "ma_map_batch_dev: null index/workspace/output");
    MA_REQUIRE(n >= 0 && n < (1ll << 32), "ma_map_batch_dev: n out of range");
    MA_REQUIRE(n <= ws->max_reads || ws->max_reads == 0, "ma_map_batch_dev: n exceeds workspace");
    MA_REQUIRE(encoding == MA_ENC_ASCII || encoding == MA_ENC_2BIT, "ma_map_batch_dev: bad encoding");
    MA_REQUIRE(encoding == MA_ENC_ASCII ? stride >= max_len : stride * 4 >= max_len,
               "ma_map_batch_dev: stride smaller than a read row");
    MA_TRY(check_params(ix, p, max_len));
    if (n == 0) return MA_OK;
    RecOut mo = rec_out(d_merged);
    MA_REQUIRE(mo.cid, "ma_map_batch_dev: merged record arrays must be non-null");
    hipStream_t st = (hipStream_t)stream;
    Knobs kn{p->k, p->stride, p->max_occ, p->vote_occ, p->vote_fast, p->max_diff, p->max_soft, p->do_rc};
    ReadsIn in{d_bases, stride, d_lens, n, max_len, encoding, bs_r1, bs_r2, n_r1};
    zero_u32_kernel<<<1, 1, 0, st>>>(ws->counter.as<unsigned int>());
    MA_LAUNCHED();
    int dev = ix->device;
    int n_cu = 0;
    MA_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    int64_t blocks = (n + WPB - 1) / WPB;
    int64_t cap = (int64_t)(n_cu > 0 ? n_cu : 256) * 8;
    if (blocks > cap) blocks = cap;
    map_reads_kernel<<<(unsigned)blocks, BLOCK, 0, st>>>(index_view(ix), kn, in, mo, rec_out(d_fw),
                                                          rec_out(d_rc), ws->counter.as<unsigned int>());
    MA_LAUNCHED();
    return MA_OK;
}

int ma_map_batch(ma_index *ix, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                 int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                 int64_t n_r1, const ma_params *p, ma_recs *merged, ma_recs *fw, ma_recs *rc) {
    MA_REQUIRE(ix && merged, "ma_map_batch: null index/output");
    MA_REQUIRE(n >= 0, "ma_map_batch: n < 0");
    MA_TRY(check_params(ix, p, max_len));
    if (n == 0) return MA_OK;
    MA_REQUIRE(bases && lens, "ma_map_batch: null input");
    for (int64_t i = 0; i < n; i++)
        MA_REQUIRE((int64_t)lens[i] <= max_len, "ma_map_batch: a read is longer than max_len");
    MA_HIP(hipSetDevice(ix->device));
    hipStream_t st = ix->stream;
    int64_t in_bytes = n * stride;
    DevBuf d_in, d_lens, d_out[3];
    MA_TRY(d_in.alloc((size_t)in_bytes));
    MA_TRY(d_lens.alloc((size_t)n * 4));
    ma_recs *host_recs[3] = {merged, fw, rc};
    ma_recs dev_recs[3];
    for (int s = 0; s < 3; s++) {
        memset(&dev_recs[s], 0, sizeof(ma_recs));
        if (!host_recs[s]) continue;
        MA_TRY(d_out[s].alloc((size_t)n * 4 * 10));
        char *b = (char *)d_out[s].p;
        dev_recs[s].cid = (int32_t *)(b + 0 * n * 4);
        dev_recs[s].pos = (uint32_t *)(b + 1 * n * 4);
        dev_recs[s].mapq = (uint32_t *)(b + 2 * n * 4);
        dev_recs[s].flag = (int32_t *)(b + 3 * n * 4);
        dev_recs[s].sl = (uint32_t *)(b + 4 * n * 4);
        dev_recs[s].sr = (uint32_t *)(b + 5 * n * 4);
        dev_recs[s].nm = (uint32_t *)(b + 6 * n * 4);
        dev_recs[s].iop = (uint32_t *)(b + 7 * n * 4);
        dev_recs[s].iat = (uint32_t *)(b + 8 * n * 4);
        dev_recs[s].ilen = (uint32_t *)(b + 9 * n * 4);
    }
    MA_HIP(hipMemcpyAsync(d_in.p, bases, (size_t)in_bytes, hipMemcpyHostToDevice, st));
    MA_HIP(hipMemcpyAsync(d_lens.p, lens, (size_t)n * 4, hipMemcpyHostToDevice, st));
    ma_workspace ws;
    ws.ix = ix;
    ws.max_reads = n;
    ws.max_len = max_len;
    MA_TRY(ws.counter.alloc(256));
    MA_TRY(ma_map_batch_dev(ix, &ws, st, d_in.as<uint8_t>(), stride, d_lens.as<uint32_t>(), n, max_len,
                            encoding, bs_r1, bs_r2, n_r1, p, &dev_recs[0],
                            host_recs[1] ? &dev_recs[1] : nullptr, host_recs[2] ? &dev_recs[2] : nullptr));
    for (int s = 0; s < 3; s++) {
        if (!host_recs[s]) continue;
        void *dst[10] = {host_recs[s]->cid, host_recs[s]->pos, host_recs[s]->mapq, host_recs[s]->flag,
                         host_recs[s]->sl,  host_recs[s]->sr,  host_recs[s]->nm,   host_recs[s]->iop,
                         host_recs[s]->iat, host_recs[s]->ilen};
        for (int f = 0; f < 10; f++) {
            if (!dst[f]) continue;
            MA_HIP(hipMemcpyAsync(dst[f], (char *)d_out[s].p + (size_t)f * n * 4, (size_t)n * 4,
                                  hipMemcpyDeviceToHost, st));
        }
    }
    MA_HIP(hipStreamSynchronize(st));
    return MA_OK;
}

int ma_seed_kmers(int device, const uint8_t *bases, int64_t n, int32_t max_len, int32_t k,
                  uint64_t *out_hash) {
    MA_REQUIRE(n >= 0 && max_len >= 1 && k >= 1 && k <= 31 && bases && out_hash,
               "ma_seed_kmers: bad arguments");
    if (n == 0) return MA_OK;
    MA_HIP(hipSetDevice(device));
    int64_t nb = n * max_len;
    DevBuf d_b, d_h;
    MA_TRY(d_b.alloc((size_t)nb));
    MA_TRY(d_h.alloc((size_t)nb * 8));
    hipStream_t st;
    MA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = MA_OK;
    if (hipMemcpyAsync(d_b.p, bases, (size_t)nb, hipMemcpyHostToDevice, st) != hipSuccess) rc = MA_ERR_HIP;
    if (rc == MA_OK) {
        seed_hash_kernel<<<grid_for(nb), 256, 0, st>>>(d_b.as<uint8_t>(), nb, max_len, k,
                                                       d_h.as<uint64_t>());
        if (hipGetLastError() != hipSuccess) rc = MA_ERR_HIP;
    }
    if (rc == MA_OK &&
        (hipMemcpyAsync(out_hash, d_h.p, (size_t)nb * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
        rc = MA_ERR_HIP;
    (void)hipStreamDestroy(st);
    if (rc != MA_OK) set_error("ma_seed_kmers: device operation failed");
    return rc;
}

}  // extern "C"
