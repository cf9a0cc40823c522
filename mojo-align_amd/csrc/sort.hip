// Sort keys (SE and PE), stable coordinate sort and Picard-style duplicate marking.
//
// Semantics follow gpu_sort_markdup (fq2bam-meth/src/linear_gpu_sort.mojo:24-551):
//   dup order  = stable sort by (dhi, dlo)   (LSD: by dlo, then by dhi)
//   groups     = runs of equal (dhi, dlo); dhi == ~0 is never a duplicate
//   best       = max score, ties -> min pair id
//   is_dup     = pair != best pair           (written in original order)
//   perm       = stable sort by coord
// Design: both orders come from the one-sweep radix sort (ma_onesweep.hpp), a
// fixed chain of kernels with no host round trip. Duplicates are marked in the
// (dhi, dlo) order: a tile's head flags are scanned into group ids (the same
// sweep clears is_dup); each wave then takes a range of consecutive sorted
// positions, gathers (score << 32 | ~pair) only for records of groups of two or
// more, finds each group's best in LDS and writes is_dup = 1 (in original
// order) for the duplicates of every group that lies inside the range; only
// groups reaching past a range meet in HBM (one 64-bit atomicMax per range)
// and are marked by a last pass over the ranges they touch.
#include <stddef.h>

#include "ma_onesweep.hpp"
#include "ma_pe.hpp"
#include "ma_sort.hpp"

using namespace ma;

namespace {

__global__ void gather_u64_kernel(const uint64_t *__restrict__ src, const uint32_t *__restrict__ idx,
                                  uint64_t *__restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

__global__ void copy_u64_kernel(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                                int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void copy_u32_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- markdup in the (dhi, dlo) order
constexpr int MD_ITEMS = 16;
constexpr int MD_TILE = 256 * MD_ITEMS;
// A wave of md_best_mark_kernel takes MD_RANGE consecutive sorted positions
// (lane l: MD_W consecutive ones).
constexpr int MD_W = 8;
constexpr int MD_RANGE = 64 * MD_W;

// The sorted dup keys and permutation as the dup sort's plan left them.
struct MdBufs {
    uint64_t *lo[2], *hi[2];
    uint32_t *v[2];
    uint32_t *out_v;  // caller's dup_perm (null: v[fin])
};

struct MdView {
    const uint64_t *hi, *lo;  // lo null when it did not move (implied by hi, or constant)
    const uint32_t *idx;      // the (dhi, dlo) order: original record of each sorted position
    uint32_t *gid;            // group of each sorted position (the free permutation buffer)
    uint64_t *best;           // per group reaching past a range: the largest (score << 32 | ~pair)
                              // of its records (the free key buffer)
};

__device__ __forceinline__ MdView md_view(const SortPlan *plan, const MdBufs &B) {
    const uint32_t f = plan->fin;
    MdView m;
    m.hi = B.hi[f];
    m.lo = plan->carry ? B.lo[f] : nullptr;
    m.idx = B.out_v ? B.out_v : B.v[f];
    m.gid = B.v[f ^ 1u];
    m.best = B.lo[f ^ 1u];
    return m;
}

__device__ __forceinline__ uint32_t md_head(const MdView &m, int64_t i) {
    if (i == 0) return 1u;
    if (m.hi[i] != m.hi[i - 1]) return 1u;
    return (m.lo && m.lo[i] != m.lo[i - 1]) ? 1u : 0u;
}

// Group heads per tile of the sorted order.
__global__ void __launch_bounds__(256) md_tile_heads_kernel(const SortPlan *__restrict__ plan, MdBufs B,
                                                             int64_t n, uint64_t *__restrict__ partial) {
    __shared__ uint64_t sh[4];
    const MdView m = md_view(plan, B);
    const int64_t base = (int64_t)blockIdx.x * MD_TILE;
    uint64_t c = 0;
    for (int r = 0; r < MD_ITEMS; r++) {
        const int64_t i = base + (int64_t)r * 256 + threadIdx.x;
        if (i < n) c += md_head(m, i);
    }
    uint64_t tot;
    (void)block_excl_scan_256(c, sh, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// gid[i] = the group of sorted position i (heads scanned within the tile on
// top of the tile's exclusive head count). The same sweep clears is_dup (only
// duplicates are written afterwards) and best[g] of every group at a range's
// first or last position (the only groups that can reach past a range).
__global__ void __launch_bounds__(256) md_gid_kernel(const SortPlan *__restrict__ plan, MdBufs B, int64_t n,
                                                      const uint64_t *__restrict__ partial,
                                                      uint32_t *__restrict__ is_dup) {
    __shared__ uint64_t sh[4];
    const MdView m = md_view(plan, B);
    const int64_t base = (int64_t)blockIdx.x * MD_TILE;
    uint64_t run = partial[blockIdx.x];
    for (int r = 0; r < MD_ITEMS; r++) {
        const int64_t r0 = base + (int64_t)r * 256;
        if (r0 >= n) break;  // block-uniform
        const int64_t i = r0 + threadIdx.x;
        const bool valid = i < n;
        const uint32_t h = valid ? md_head(m, i) : 0u;
        uint64_t tot;
        const uint64_t ex = block_excl_scan_256(h, sh, tot);
        if (valid) {
            const uint32_t g = (uint32_t)(run + ex + h - 1);
            m.gid[i] = g;
            is_dup[i] = 0u;
            const int64_t rr = i % MD_RANGE;
            if (rr == 0 || rr == MD_RANGE - 1) m.best[g] = 0ull;
        }
        run += tot;
    }
}

// Best + mark, fused. A wave takes ranges of MD_RANGE consecutive sorted
// positions. A record alone in its group, or unmapped (dhi == ~0), is never a
// duplicate and costs nothing past the sorted arrays' sweep. A record of a
// larger group gathers its (score << 32 | ~pair) once and folds it into its
// group's slot in LDS: the group ids of a range are consecutive, so group g
// has slot g - (the range's first id), one of at most MD_RANGE. A group whose
// positions all lie in the range is resolved at once: a record is a duplicate
// unless its pair is the group's best, and only duplicates are written. The
// range's first and last group, when they reach past the range, fold their
// part into best[g] in HBM instead (one atomic each) and are marked by
// md_mark_edges_kernel, so a group of any size costs at most one atomic per
// range it spans.
__device__ __forceinline__ bool md_first_spills(const MdView &m, int64_t b) {
    return b > 0 && m.gid[b - 1] == m.gid[b];
}

__device__ __forceinline__ bool md_last_spills(const MdView &m, int64_t e, int64_t n) {
    return e < n && m.gid[e] == m.gid[e - 1];
}

constexpr uint32_t MD_NONE = 0xFFFFFFFFu;  // no group (ids are < n <= 2^32 - 1)

__global__ void __launch_bounds__(256) md_best_mark_kernel(const SortPlan *__restrict__ plan, MdBufs B, int64_t n,
                                                           const uint32_t *__restrict__ score,
                                                           const uint32_t *__restrict__ pair,
                                                           uint32_t *__restrict__ is_dup) {
    __shared__ unsigned long long s_best[4][MD_RANGE];
    const MdView m = md_view(plan, B);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long *sb = s_best[w];
    for (int64_t b = ((int64_t)blockIdx.x * 4 + w) * MD_RANGE; b < n; b += (int64_t)gridDim.x * 4 * MD_RANGE) {
        const int64_t e = b + MD_RANGE < n ? b + MD_RANGE : n;
        const uint32_t g0 = m.gid[b];
        const uint32_t g_first_spill = md_first_spills(m, b) ? g0 : MD_NONE;
        const uint32_t g_last_spill = md_last_spills(m, e, n) ? m.gid[e - 1] : MD_NONE;
        const int64_t i0 = b + (int64_t)lane * MD_W;
        uint32_t o[MD_W], g[MD_W];
        bool mapped[MD_W];
#pragma unroll
        for (int j = 0; j < MD_W; j++) {
            const int64_t i = i0 + j;
            o[j] = MD_NONE;
            g[j] = MD_NONE;
            mapped[j] = false;
            if (i < e) {
                o[j] = m.idx[i];
                g[j] = m.gid[i];
                mapped[j] = m.hi[i] != ~0ull;
            }
        }
        // the neighbours of the lane's first and last position (positions past
        // e on this lane only occur at the end of the sorted order)
        const uint32_t g_prev = (i0 < e && i0 > 0) ? m.gid[i0 - 1] : MD_NONE;
        const uint32_t g_next = (i0 + MD_W - 1 < e && i0 + MD_W < n) ? m.gid[i0 + MD_W] : MD_NONE;
        bool need[MD_W];
#pragma unroll
        for (int j = 0; j < MD_W; j++) {
            const uint32_t gp = j > 0 ? g[j - 1] : g_prev, gn = j < MD_W - 1 ? g[j + 1] : g_next;
            // o >= n only after a radix pass's timeout (plan->error)
            need[j] = mapped[j] && (int64_t)o[j] < n && (gp == g[j] || gn == g[j]);
        }
        const uint64_t any = __ballot(need[0] || need[1] || need[2] || need[3] || need[4] || need[5] ||
                                      need[6] || need[7]);
        if (any == 0ull) continue;  // wave-uniform: no group of two or more in the range
#pragma unroll
        for (int j = 0; j < MD_W; j++) sb[lane * MD_W + j] = 0ull;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t xs[MD_W], xp[MD_W];
#pragma unroll
        for (int j = 0; j < MD_W; j++) {
            xs[j] = need[j] ? score[o[j]] : 0u;
            xp[j] = need[j] ? 0xFFFFFFFFu - pair[o[j]] : 0u;
        }
#pragma unroll
        for (int j = 0; j < MD_W; j++)
            if (need[j]) atomicMax(&sb[g[j] - g0], ((unsigned long long)xs[j] << 32) | xp[j]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the spilling groups' parts go to HBM (their first position in the range)
#pragma unroll
        for (int j = 0; j < MD_W; j++) {
            if (!need[j]) continue;
            const bool spill = g[j] == g_first_spill || g[j] == g_last_spill;
            const unsigned long long best = sb[g[j] - g0];
            if (spill) {
                const int64_t i = i0 + j;
                const uint32_t gp = j > 0 ? g[j - 1] : g_prev;
                if (i == b || gp != g[j]) atomicMax((unsigned long long *)&m.best[g[j]], best);
            } else if (xp[j] != (uint32_t)best) {
                is_dup[o[j]] = 1u;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// The records of the groups that reach past a range, marked against best[g]:
// a wave per range, only for ranges whose first or last group spills.
__global__ void __launch_bounds__(256) md_mark_edges_kernel(const SortPlan *__restrict__ plan, MdBufs B, int64_t n,
                                                            const uint32_t *__restrict__ pair,
                                                            uint32_t *__restrict__ is_dup) {
    const MdView m = md_view(plan, B);
    const int lane = threadIdx.x & 63;
    const int64_t nr = (n + MD_RANGE - 1) / MD_RANGE;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < nr; r += nw) {
        const int64_t b = r * MD_RANGE;
        const int64_t e = b + MD_RANGE < n ? b + MD_RANGE : n;
        const uint32_t fs = md_first_spills(m, b) ? m.gid[b] : MD_NONE;
        const uint32_t ls = md_last_spills(m, e, n) ? m.gid[e - 1] : MD_NONE;
        if (fs == MD_NONE && ls == MD_NONE) continue;  // wave-uniform
        for (int64_t i = b + lane; i < e; i += 64) {
            const uint32_t gi = m.gid[i];
            if (gi != fs && gi != ls) continue;
            const uint32_t o = m.idx[i];
            if ((int64_t)o >= n || m.hi[i] == ~0ull) continue;
            if (0xFFFFFFFFu - pair[o] != (uint32_t)m.best[gi]) is_dup[o] = 1u;
        }
    }
}

// _unclipped5 (linear_gpu_fm.mojo:371-387): the 5' end on the reference,
// clips counted in; 0 for an unmapped record or an empty read.
__device__ __forceinline__ int64_t unclipped5(int64_t fl, int64_t p0, int64_t a, int64_t b, int64_t ql) {
    if ((fl & 4) != 0 || ql <= 0) return 0;
    if (a + b >= ql) a = b = 0;
    const int64_t mid = ql - a - b;
    return (fl & 16) ? p0 + mid + a - 1 : p0 - a;
}

// _pack_end (:398-405): (tid + 1) << 48 | strand << 47 | (u5 + 2^30) in 47 bits.
__device__ __forceinline__ uint64_t pack_end(int64_t tid, int64_t u5, int64_t strand) {
    int64_t u = u5 + 1073741824;
    if (u < 0) u = 0;
    return ((uint64_t)(tid >= 0 ? tid + 1 : 0) << 48) | ((uint64_t)(strand & 1) << 47) |
           ((uint64_t)u & 140737488355327ull);
}

// _coord_key (:390-396): (tid << 32 | pos), unmapped (0xFFFFFFFF << 32 | orig).
__device__ __forceinline__ uint64_t coord_key(int64_t tid, int64_t p0, int64_t orig) {
    if (tid < 0) return (0xFFFFFFFFull << 32) | ((uint64_t)orig & 0xFFFFFFFFull);
    return ((uint64_t)tid << 32) | (uint64_t)(p0 < 0 ? 0 : p0);
}

// SE sort keys; follows the SE builder in linear_gpu_fm.mojo:800-823.
__global__ void sort_keys_se_kernel(int64_t n, const int32_t *__restrict__ cid,
                                    const uint32_t *__restrict__ pos, const int32_t *__restrict__ flag,
                                    const uint32_t *__restrict__ sl, const uint32_t *__restrict__ sr,
                                    const uint32_t *__restrict__ nm, const uint32_t *__restrict__ mapq,
                                    const uint32_t *__restrict__ lens,
                                    const int32_t *__restrict__ tid_of_cid, int64_t n_tid, int64_t pair_base,
                                    uint64_t *__restrict__ coord, uint64_t *__restrict__ dhi,
                                    uint64_t *__restrict__ dlo, uint32_t *__restrict__ score,
                                    uint32_t *__restrict__ pairo) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = cid[i];
        // a contig outside the target table has no target (rid1 < len(rid_to_tid),
        // linear_gpu_fm.mojo:764-765)
        int64_t tid = c >= 0 && c < n_tid ? (int64_t)tid_of_cid[c] : -1;
        int64_t fl = flag[i];
        int64_t p0 = (int64_t)pos[i];
        coord[i] = coord_key(tid, p0, i);
        if (tid < 0) {
            dhi[i] = ~0ull;
            dlo[i] = (uint64_t)i;
        } else {
            dhi[i] = pack_end(tid, unclipped5(fl, p0, sl[i], sr[i], lens[i]), (fl >> 4) & 1);
            dlo[i] = 0;
        }
        uint32_t nmc = nm[i] > 255u ? 255u : nm[i];
        score[i] = (mapq[i] << 8) | (255u - nmc);
        pairo[i] = (uint32_t)(pair_base + i);
    }
}

// PE sort keys, a thread per pair writing output records 2p (R1) and 2p + 1
// (R2); follows the PE builder in linear_gpu_fm.mojo:936-977: coord per record,
// the pair's two ends ordered into (dhi, dlo) (one end when a mate has no
// target; ~0 and the R1 index when neither has), score (mapq1 + mapq2) << 8 |
// 255 - min(nm1 + nm2, 255) and pair id (pair_base + 2p) / 2 on both records.
__global__ void sort_keys_pe_kernel(int64_t np, const int32_t *__restrict__ cid,
                                    const uint32_t *__restrict__ pos, const int32_t *__restrict__ flag,
                                    const uint32_t *__restrict__ sl, const uint32_t *__restrict__ sr,
                                    const uint32_t *__restrict__ nm, const uint32_t *__restrict__ mapq,
                                    const uint32_t *__restrict__ lens,
                                    const int32_t *__restrict__ tid_of_cid, int64_t n_tid, int64_t pair_base,
                                    uint64_t *__restrict__ coord, uint64_t *__restrict__ dhi,
                                    uint64_t *__restrict__ dlo, uint32_t *__restrict__ score,
                                    uint32_t *__restrict__ pairo) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np;
         p += (int64_t)gridDim.x * blockDim.x) {
        PeMate m[2];
        pe_pair(np, p, cid, pos, flag, lens, tid_of_cid, n_tid, m);
        const int64_t r1 = p, r2 = np + p;
        uint64_t h = ~0ull, l = (uint64_t)(2 * p);
        if (m[0].tid >= 0 || m[1].tid >= 0) {
            const uint64_t e1 = pack_end(m[0].tid, unclipped5(m[0].flag, m[0].pos, sl[r1], sr[r1], lens[r1]),
                                         (m[0].flag >> 4) & 1);
            const uint64_t e2 = pack_end(m[1].tid, unclipped5(m[1].flag, m[1].pos, sl[r2], sr[r2], lens[r2]),
                                         (m[1].flag >> 4) & 1);
            if (m[0].tid < 0) {
                h = e2;
                l = 0;
            } else if (m[1].tid < 0) {
                h = e1;
                l = 0;
            } else {
                h = e1 <= e2 ? e1 : e2;
                l = e1 <= e2 ? e2 : e1;
            }
        }
        const uint64_t nms = (uint64_t)nm[r1] + nm[r2];
        const uint32_t sc =
            (uint32_t)((((uint64_t)mapq[r1] + mapq[r2]) << 8) | (255u - (nms > 255 ? 255u : (uint32_t)nms)));
        const uint32_t pid = (uint32_t)((pair_base + 2 * p) / 2);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int64_t i = 2 * p + k;
            coord[i] = coord_key(m[k].tid, m[k].pos, i);
            dhi[i] = h;
            dlo[i] = l;
            score[i] = sc;
            pairo[i] = pid;
        }
    }
}

struct SortLayout {
    size_t idx_a, idx_b, key_a, key_b, heads, excl, best, partial, bh, bo, h8, total;
};

SortLayout sort_layout(int64_t n) {
    SortLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    int64_t nb = rs_nblocks(n);
    L.idx_a = take((size_t)n * 4);
    L.idx_b = take((size_t)n * 4);
    L.key_a = take((size_t)n * 8);
    L.key_b = take((size_t)n * 8);
    L.heads = take((size_t)n * 4);
    L.excl = take((size_t)n * 8);
    L.best = take((size_t)n * 8);
    L.partial = take(scan_scratch_bytes(n));
    L.bh = take((size_t)(256 * nb * 4));
    L.bo = take((size_t)(256 * nb * 8));
    L.h8 = take(8 * 256 * 8);
    L.total = off;
    return L;
}

// Scratch of ma_sort_markdup_dev: the one-sweep buffers (two of each key word
// and of the permutation), the look-back status arrays, histograms, the two
// sorts' plans, tile counters and the markdup tile partials.
struct MdLayout {
    size_t lo0, lo1, hi0, hi1, v0, v1, status, hist, plan_dup, plan_coord, tile_ctr, partial, total;
    int64_t status_stride;
};

MdLayout md_layout(int64_t n) {
    MdLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t nn = (size_t)(n < 1 ? 1 : n);
    L.lo0 = take(nn * 8);
    L.lo1 = take(nn * 8);
    L.hi0 = take(nn * 8);
    L.hi1 = take(nn * 8);
    L.v0 = take(nn * 4);
    L.v1 = take(nn * 4);
    L.status_stride = os_status_entries(n);
    L.status = take((size_t)L.status_stride * 2 * 8);
    L.hist = take(OS_PASSES * OS_RADIX * 8);
    L.plan_dup = take(sizeof(SortPlan));
    L.plan_coord = take(sizeof(SortPlan));
    L.tile_ctr = take(OS_PASSES * OS_GROUPS * 4);
    L.partial = take((size_t)((n + MD_TILE - 1) / MD_TILE + 1) * 8);
    L.total = off;
    return L;
}

OsScratch os_scratch(char *s, const MdLayout &L, size_t plan_off) {
    OsScratch S;
    S.lo[0] = (uint64_t *)(s + L.lo0);
    S.lo[1] = (uint64_t *)(s + L.lo1);
    S.hi[0] = (uint64_t *)(s + L.hi0);
    S.hi[1] = (uint64_t *)(s + L.hi1);
    S.v[0] = (uint32_t *)(s + L.v0);
    S.v[1] = (uint32_t *)(s + L.v1);
    S.status = (uint64_t *)(s + L.status);
    S.status_stride = L.status_stride;
    S.hist = (unsigned long long *)(s + L.hist);
    S.plan = (SortPlan *)(s + plan_off);
    S.tile_ctr = (uint32_t *)(s + L.tile_ctr);
    return S;
}

}  // namespace

namespace ma {
void sort_error_words(int64_t n, size_t *dup_word, size_t *coord_word) {
    const MdLayout L = md_layout(n < 1 ? 1 : n);
    *dup_word = L.plan_dup + offsetof(SortPlan, error);
    *coord_word = L.plan_coord + offsetof(SortPlan, error);
}
}  // namespace ma

extern "C" {

int64_t ma_sort_scratch_bytes(int64_t n) {
    const size_t a = sort_layout(n < 1 ? 1 : n).total, b = md_layout(n < 1 ? 1 : n).total;
    return (int64_t)(a > b ? a : b);
}

int ma_radix_sort_u64_dev(void *stream, int64_t n, uint64_t *d_keys, uint32_t *d_vals,
                          void *d_scratch, int64_t scratch_bytes) {
    MA_REQUIRE(n >= 0, "ma_radix_sort_u64_dev: n < 0");
    if (n <= 1) return MA_OK;
    SortLayout L = sort_layout(n);
    MA_REQUIRE(scratch_bytes >= (int64_t)L.total, "ma_radix_sort_u64_dev: scratch too small");
    hipStream_t st = (hipStream_t)stream;
    char *s = (char *)d_scratch;
    RadixScratch rs{(uint32_t *)(s + L.bh), (uint64_t *)(s + L.bo), (unsigned long long *)(s + L.h8)};
    uint64_t *k = d_keys, *ka = (uint64_t *)(s + L.key_a);
    uint32_t *v = d_vals, *va = (uint32_t *)(s + L.idx_a);
    MA_TRY(radix_sort_pairs<uint32_t>(st, n, k, v, ka, va, rs));
    if (k != d_keys) {
        copy_u64_kernel<<<grid_for(n), 256, 0, st>>>(k, d_keys, n);
        MA_LAUNCHED();
        copy_u32_kernel<<<grid_for(n), 256, 0, st>>>(v, d_vals, n);
        MA_LAUNCHED();
    }
    return MA_OK;
}

int ma_sort_keys_se_dev(void *stream, int64_t n, const ma_recs *m, const uint32_t *d_lens,
                        const int32_t *d_tid_of_cid, int32_t n_tid, int64_t pair_base, uint64_t *d_coord,
                        uint64_t *d_dhi, uint64_t *d_dlo, uint32_t *d_score, uint32_t *d_pair) {
    MA_REQUIRE(n >= 0 && m && n_tid >= 0, "ma_sort_keys_se_dev: bad arguments");
    if (n == 0) return MA_OK;
    sort_keys_se_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        n, m->cid, m->pos, m->flag, m->sl, m->sr, m->nm, m->mapq, d_lens, d_tid_of_cid, n_tid, pair_base,
        d_coord, d_dhi, d_dlo, d_score, d_pair);
    MA_LAUNCHED();
    return MA_OK;
}

int ma_sort_keys_pe_dev(void *stream, int64_t n_pairs, const ma_recs *m, const uint32_t *d_lens,
                        const int32_t *d_tid_of_cid, int32_t n_tid, int64_t pair_base, uint64_t *d_coord,
                        uint64_t *d_dhi, uint64_t *d_dlo, uint32_t *d_score, uint32_t *d_pair) {
    MA_REQUIRE(n_pairs >= 0 && m && n_tid >= 0, "ma_sort_keys_pe_dev: bad arguments");
    MA_REQUIRE(2 * n_pairs <= 0xFFFFFFFFll, "ma_sort_keys_pe_dev: records exceed u32 ids");
    if (n_pairs == 0) return MA_OK;
    sort_keys_pe_kernel<<<grid_for(n_pairs), 256, 0, (hipStream_t)stream>>>(
        n_pairs, m->cid, m->pos, m->flag, m->sl, m->sr, m->nm, m->mapq, d_lens, d_tid_of_cid, n_tid, pair_base,
        d_coord, d_dhi, d_dlo, d_score, d_pair);
    MA_LAUNCHED();
    return MA_OK;
}

int ma_sort_markdup_dev(void *stream, int64_t n, const uint64_t *d_coord, const uint64_t *d_dhi,
                        const uint64_t *d_dlo, const uint32_t *d_score, const uint32_t *d_pair,
                        uint32_t *d_is_dup, uint32_t *d_perm, int32_t do_markdup,
                        uint32_t *d_dup_perm, void *d_scratch, int64_t scratch_bytes) {
    MA_REQUIRE(n >= 0, "ma_sort_markdup_dev: n < 0");
    MA_REQUIRE(n <= 0xFFFFFFFFll, "ma_sort_markdup_dev: n exceeds u32 record ids");
    if (n == 0) return MA_OK;
    MA_REQUIRE(scratch_bytes >= ma_sort_scratch_bytes(n), "ma_sort_markdup_dev: scratch too small");
    hipStream_t st = (hipStream_t)stream;
    const MdLayout L = md_layout(n);
    char *s = (char *)d_scratch;
    const OsScratch S = os_scratch(s, L, L.plan_dup);
    const int g = grid_for(n);
    if (do_markdup) {
        MA_TRY(os_sort(st, n, d_dlo, d_dhi, d_dup_perm, 1, 1, S));
        const MdBufs B{{S.lo[0], S.lo[1]}, {S.hi[0], S.hi[1]}, {S.v[0], S.v[1]}, d_dup_perm};
        const int64_t nt = (n + MD_TILE - 1) / MD_TILE;
        uint64_t *partial = (uint64_t *)(s + L.partial);
        md_tile_heads_kernel<<<(unsigned)nt, 256, 0, st>>>(S.plan, B, n, partial);
        MA_LAUNCHED();
        scan_top_kernel<<<1, RS_BLOCK, 0, st>>>(partial, nt, nullptr);
        MA_LAUNCHED();
        md_gid_kernel<<<(unsigned)nt, 256, 0, st>>>(S.plan, B, n, partial, d_is_dup);
        MA_LAUNCHED();
        const int64_t nr = (n + MD_RANGE - 1) / MD_RANGE;
        const int gr = (int)((nr + 3) / 4 < (1 << 20) ? (nr + 3) / 4 : (1 << 20));
        md_best_mark_kernel<<<gr, 256, 0, st>>>(S.plan, B, n, d_score, d_pair, d_is_dup);
        MA_LAUNCHED();
        md_mark_edges_kernel<<<gr, 256, 0, st>>>(S.plan, B, n, d_pair, d_is_dup);
        MA_LAUNCHED();
    } else {
        fill_u32_kernel<<<g, 256, 0, st>>>(d_is_dup, n, 0u);
        MA_LAUNCHED();
        // no duplicate-order sort runs: clear its error word, which
        // ma_sort_dev_status reads whatever the scratch held before
        fill_u32_kernel<<<1, 64, 0, st>>>((uint32_t *)(s + L.plan_dup + offsetof(SortPlan, error)), 1, 0u);
        MA_LAUNCHED();
        if (d_dup_perm) {
            iota_u32_kernel<<<g, 256, 0, st>>>(d_dup_perm, n);
            MA_LAUNCHED();
        }
    }
    const OsScratch S2 = os_scratch(s, L, L.plan_coord);
    MA_TRY(os_sort(st, n, d_coord, nullptr, d_perm, 0, 0, S2));
    return MA_OK;
}

int ma_sort_dev_status(void *stream, int64_t n, const void *d_scratch, int64_t scratch_bytes) {
    MA_REQUIRE(n >= 0 && d_scratch && scratch_bytes >= ma_sort_scratch_bytes(n), "ma_sort_dev_status: bad arguments");
    if (n == 0) return MA_OK;
    const MdLayout L = md_layout(n);
    uint32_t err[2] = {0, 0};
    hipStream_t st = (hipStream_t)stream;
    MA_HIP(hipMemcpyAsync(&err[0], (const char *)d_scratch + L.plan_dup + offsetof(SortPlan, error), 4,
                          hipMemcpyDeviceToHost, st));
    MA_HIP(hipMemcpyAsync(&err[1], (const char *)d_scratch + L.plan_coord + offsetof(SortPlan, error), 4,
                          hipMemcpyDeviceToHost, st));
    MA_HIP(hipStreamSynchronize(st));
    if (err[0] || err[1]) {
        set_error("ma_sort_markdup: a radix pass timed out waiting for a preceding tile (results invalid)");
        return MA_ERR_HIP;
    }
    return MA_OK;
}

int ma_sort_markdup(int device, int64_t n, const uint64_t *coord, const uint64_t *dhi,
                    const uint64_t *dlo, const uint32_t *score, const uint32_t *pair,
                    uint32_t *is_dup, uint32_t *perm, int32_t do_markdup, uint32_t *dup_perm) {
    MA_REQUIRE(n >= 0, "ma_sort_markdup: n < 0");
    if (n == 0) return MA_OK;
    MA_REQUIRE(coord && dhi && dlo && score && pair && is_dup && perm,
               "ma_sort_markdup: null array");
    MA_HIP(hipSetDevice(device));
    DevBuf b_coord, b_dhi, b_dlo, b_score, b_pair, b_dup, b_perm, b_dperm, b_scr;
    MA_TRY(b_coord.alloc((size_t)n * 8));
    MA_TRY(b_dhi.alloc((size_t)n * 8));
    MA_TRY(b_dlo.alloc((size_t)n * 8));
    MA_TRY(b_score.alloc((size_t)n * 4));
    MA_TRY(b_pair.alloc((size_t)n * 4));
    MA_TRY(b_dup.alloc((size_t)n * 4));
    MA_TRY(b_perm.alloc((size_t)n * 4));
    MA_TRY(b_dperm.alloc((size_t)n * 4));
    int64_t sb = ma_sort_scratch_bytes(n);
    MA_TRY(b_scr.alloc((size_t)sb));
    hipStream_t st;
    MA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = MA_OK;
    do {
        if (hipMemcpyAsync(b_coord.p, coord, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(b_dhi.p, dhi, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(b_dlo.p, dlo, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(b_score.p, score, (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(b_pair.p, pair, (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
            set_error("ma_sort_markdup: host->device copy failed");
            rc = MA_ERR_HIP;
            break;
        }
        rc = ma_sort_markdup_dev(st, n, b_coord.as<uint64_t>(), b_dhi.as<uint64_t>(),
                                 b_dlo.as<uint64_t>(), b_score.as<uint32_t>(), b_pair.as<uint32_t>(),
                                 b_dup.as<uint32_t>(), b_perm.as<uint32_t>(), do_markdup,
                                 b_dperm.as<uint32_t>(), b_scr.p, sb);
        if (rc != MA_OK) break;
        rc = ma_sort_dev_status(st, n, b_scr.p, sb);
        if (rc != MA_OK) break;
        if (hipMemcpyAsync(is_dup, b_dup.p, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(perm, b_perm.p, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            (dup_perm && hipMemcpyAsync(dup_perm, b_dperm.p, (size_t)n * 4, hipMemcpyDeviceToHost,
                                        st) != hipSuccess) ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("ma_sort_markdup: device->host copy failed");
            rc = MA_ERR_HIP;
        }
    } while (0);
    (void)hipStreamDestroy(st);
    return rc;
}

}  // extern "C"
