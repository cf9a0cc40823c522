// Sort keys (SE and PE), stable coordinate sort and Picard-style duplicate marking.
//
// Semantics follow gpu_sort_markdup (fq2bam-meth/src/linear_gpu_sort.mojo:24-551):
//   dup order  = stable sort by (dhi, dlo)   (LSD: by dlo, then by dhi)
//   groups     = runs of equal (dhi, dlo); dhi == ~0 is never a duplicate
//   best       = max score, ties -> min pair id
//   is_dup     = pair != best pair           (written in original order)
//   perm       = stable sort by coord
// Design: everything stays on the device; the group best is one 64-bit
// atomicMax per record on (score << 32 | ~pair) keyed by the group index from
// an exclusive scan of head flags (no host scan, no host merge).
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

// heads[i] = record i of the (dhi, dlo) order starts a new run.
__global__ void md_heads_kernel(const uint64_t *__restrict__ dhi_s, const uint32_t *__restrict__ idx,
                                const uint64_t *__restrict__ dlo, uint32_t *__restrict__ heads,
                                int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t h = 1;
        if (i > 0 && dhi_s[i] == dhi_s[i - 1] && dlo[idx[i]] == dlo[idx[i - 1]]) h = 0;
        heads[i] = h;
    }
}

__global__ void md_best_kernel(const uint64_t *__restrict__ dhi_s, const uint32_t *__restrict__ idx,
                               const uint32_t *__restrict__ heads, const uint64_t *__restrict__ excl,
                               const uint32_t *__restrict__ score, const uint32_t *__restrict__ pair,
                               unsigned long long *__restrict__ best, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (dhi_s[i] == ~0ull) continue;
        uint64_t seg = excl[i] + heads[i] - 1;
        uint32_t o = idx[i];
        unsigned long long packed =
            ((unsigned long long)score[o] << 32) | (unsigned long long)(0xFFFFFFFFu - pair[o]);
        atomicMax(&best[seg], packed);
    }
}

__global__ void md_mark_kernel(const uint64_t *__restrict__ dhi_s, const uint32_t *__restrict__ idx,
                               const uint32_t *__restrict__ heads, const uint64_t *__restrict__ excl,
                               const uint32_t *__restrict__ pair,
                               const unsigned long long *__restrict__ best,
                               uint32_t *__restrict__ is_dup, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t o = idx[i];
        uint32_t d = 0;
        if (dhi_s[i] != ~0ull) {
            uint64_t seg = excl[i] + heads[i] - 1;
            uint32_t best_pair = 0xFFFFFFFFu - (uint32_t)(best[seg] & 0xFFFFFFFFull);
            d = pair[o] != best_pair ? 1u : 0u;
        }
        is_dup[o] = d;
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

}  // namespace

extern "C" {

int64_t ma_sort_scratch_bytes(int64_t n) { return (int64_t)sort_layout(n < 1 ? 1 : n).total; }

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
    SortLayout L = sort_layout(n);
    MA_REQUIRE(scratch_bytes >= (int64_t)L.total, "ma_sort_markdup_dev: scratch too small");
    hipStream_t st = (hipStream_t)stream;
    char *s = (char *)d_scratch;
    RadixScratch rs{(uint32_t *)(s + L.bh), (uint64_t *)(s + L.bo), (unsigned long long *)(s + L.h8)};
    uint32_t *I0 = (uint32_t *)(s + L.idx_a), *I1 = (uint32_t *)(s + L.idx_b);
    uint64_t *K0 = (uint64_t *)(s + L.key_a), *K1 = (uint64_t *)(s + L.key_b);
    uint32_t *heads = (uint32_t *)(s + L.heads);
    uint64_t *excl = (uint64_t *)(s + L.excl);
    unsigned long long *best = (unsigned long long *)(s + L.best);
    uint64_t *partial = (uint64_t *)(s + L.partial);
    int g = grid_for(n);
    if (do_markdup) {
        iota_u32_kernel<<<g, 256, 0, st>>>(I0, n);
        MA_LAUNCHED();
        copy_u64_kernel<<<g, 256, 0, st>>>(d_dlo, K0, n);
        MA_LAUNCHED();
        uint64_t *kp = K0, *ka = K1;
        uint32_t *ip = I0, *ia = I1;
        MA_TRY(radix_sort_pairs<uint32_t>(st, n, kp, ip, ka, ia, rs));
        gather_u64_kernel<<<g, 256, 0, st>>>(d_dhi, ip, ka, n);
        MA_LAUNCHED();
        uint64_t *kd = ka, *kda = kp;
        uint32_t *id = ip, *ida = ia;
        MA_TRY(radix_sort_pairs<uint32_t>(st, n, kd, id, kda, ida, rs));
        md_heads_kernel<<<g, 256, 0, st>>>(kd, id, d_dlo, heads, n);
        MA_LAUNCHED();
        MA_TRY(exclusive_scan_u32(st, n, heads, excl, partial, nullptr));
        MA_TRY(fill_u64(st, (uint64_t *)best, n, 0));
        md_best_kernel<<<g, 256, 0, st>>>(kd, id, heads, excl, d_score, d_pair, best, n);
        MA_LAUNCHED();
        md_mark_kernel<<<g, 256, 0, st>>>(kd, id, heads, excl, d_pair, best, d_is_dup, n);
        MA_LAUNCHED();
        if (d_dup_perm) {
            copy_u32_kernel<<<g, 256, 0, st>>>(id, d_dup_perm, n);
            MA_LAUNCHED();
        }
    } else {
        fill_u32_kernel<<<g, 256, 0, st>>>(d_is_dup, n, 0u);
        MA_LAUNCHED();
        if (d_dup_perm) {
            iota_u32_kernel<<<g, 256, 0, st>>>(d_dup_perm, n);
            MA_LAUNCHED();
        }
    }
    iota_u32_kernel<<<g, 256, 0, st>>>(I0, n);
    MA_LAUNCHED();
    copy_u64_kernel<<<g, 256, 0, st>>>(d_coord, K0, n);
    MA_LAUNCHED();
    uint64_t *kc = K0, *kca = K1;
    uint32_t *ic = I0, *ica = I1;
    MA_TRY(radix_sort_pairs<uint32_t>(st, n, kc, ic, kca, ica, rs));
    copy_u32_kernel<<<g, 256, 0, st>>>(ic, d_perm, n);
    MA_LAUNCHED();
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
