// Resident index: 2-bit packed reference + N flags, dense-v1 arrays
// (kmers / offsets / postings) and a bucketed hash table over the k-mer keys.
//
// The dense-v1 layout is the reference's on-disk format
// (fq2bam-meth/python/mojo_linear_pack.py:1-11, 208-287): sorted unique
// MSB-first 2-bit k-mers, u64 CSR offsets, u32 (contig, pos) postings ordered
// by (key, contig, pos). The build here runs on the GPU: one key per window,
// a stable radix sort by key, head flags + scan for the CSR. Lookups do not
// binary-search kmers (locate_bsearch_kernel, linear_gpu_locate.mojo:453-490);
// they load a key's two 64-byte buckets of a two-choice hash table that maps
// key -> (start, occ), so the resulting [start, end) ranges are the same.
#include <fcntl.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "ma_index.hpp"
#include "ma_sort.hpp"

using namespace ma;

namespace {

// ASCII -> 2-bit (A0 C1 G2 T3, anything else A + N flag), one word per thread.
__global__ void pack_ref_kernel(const uint8_t *__restrict__ seq, int64_t n_bases,
                                uint64_t *__restrict__ ref2, uint64_t *__restrict__ refN,
                                int64_t n_words, int32_t *__restrict__ has_n) {
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n_words;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint64_t a = 0, nf = 0;
        int64_t b0 = w * 32;
#pragma unroll 8
        for (int j = 0; j < 32; j++) {
            int64_t g = b0 + j;
            uint64_t c = 0, isn = 0;
            if (g < n_bases) {
                uint8_t b = seq[g];
                switch (b) {
                case 'A': case 'a': c = 0; break;
                case 'C': case 'c': c = 1; break;
                case 'G': case 'g': c = 2; break;
                case 'T': case 't': c = 3; break;
                default: isn = 1;
                }
            }
            a |= c << (62 - 2 * j);
            nf |= isn << (62 - 2 * j);
        }
        ref2[w] = a;
        refN[w] = nf;
        if (nf) atomicOr(has_n, 1);
    }
}

__global__ void key_heads_kernel(const uint64_t *__restrict__ keys, int64_t n,
                                 uint32_t *__restrict__ heads) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void csr_kernel(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ locs,
                           const uint32_t *__restrict__ heads, const uint64_t *__restrict__ excl,
                           int64_t n, uint64_t *__restrict__ kmers, uint64_t *__restrict__ offsets,
                           uint32_t *__restrict__ postings) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (heads[i]) {
            kmers[excl[i]] = keys[i];
            offsets[excl[i]] = (uint64_t)i;
        }
        uint64_t l = locs[i];
        postings[2 * i] = (uint32_t)(l >> 32);
        postings[2 * i + 1] = (uint32_t)(l & 0xFFFFFFFFull);
    }
}

__global__ void set_u64_kernel(uint64_t *p, uint64_t v) { *p = v; }

// ------------------------------------------------ key-range partitioned build
// The build walks the reference's windows several times, each time keeping
// only the windows whose key lies in one key range (a "chunk"), so no more
// than one chunk's (key, location) pairs and sort buffers are ever resident.
// Chunks are unions of consecutive 2^bits key-prefix partitions sized from an
// exact histogram. Within a chunk, windows are written in position order (per
// block counts, a scan, then ordered writes), so the stable key sort yields
// postings ordered (key, contig, pos) exactly as a whole-reference sort does.
constexpr int PART_ITEMS = 16;                    // consecutive positions per thread
constexpr int64_t PART_TILE = 256 * PART_ITEMS;   // positions per block

struct WindowScan {
    const uint64_t *ref2, *refN, *coff;
    int64_t n_contigs, n_bases;
    int k;
};

// Key of the window at g (~0 when it leaves its contig or holds an N); c is a
// contig cursor that only moves forward.
__device__ inline uint64_t window_key_at(const WindowScan &w, int64_t g, int64_t &c, int64_t *loc) {
    while (c + 1 < w.n_contigs && (int64_t)w.coff[c + 1] <= g) c++;
    const int64_t p = g - (int64_t)w.coff[c];
    const int64_t clen = (int64_t)(w.coff[c + 1] - w.coff[c]);
    *loc = (c << 32) | p;
    if (p + w.k > clen) return ~0ull;
    if ((window32(w.refN, g) >> (64 - 2 * w.k)) != 0) return ~0ull;
    return window32(w.ref2, g) >> (64 - 2 * w.k);
}

__device__ inline int64_t first_contig(const WindowScan &w, int64_t g) {
    int64_t lo = 0, hi = w.n_contigs;  // last contig whose start <= g
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)w.coff[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Histogram of valid windows by key >> shift (nbins <= 1024).
__global__ void __launch_bounds__(256) part_hist_kernel(WindowScan w, int shift, int nbins,
                                                        unsigned long long *__restrict__ hist) {
    __shared__ unsigned int h[1024];
    for (int i = threadIdx.x; i < nbins; i += 256) h[i] = 0;
    __syncthreads();
    const int64_t g0 = (int64_t)blockIdx.x * PART_TILE + (int64_t)threadIdx.x * PART_ITEMS;
    if (g0 < w.n_bases) {
        int64_t c = first_contig(w, g0), loc;
        for (int j = 0; j < PART_ITEMS; j++) {
            const int64_t g = g0 + j;
            if (g >= w.n_bases) break;
            const uint64_t key = window_key_at(w, g, c, &loc);
            if (key != ~0ull) atomicAdd(&h[key >> shift], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += 256)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// Per block: windows with key in [klo, khi).
__global__ void __launch_bounds__(256) part_count_kernel(WindowScan w, uint64_t klo, uint64_t khi,
                                                         uint32_t *__restrict__ counts) {
    __shared__ uint64_t sh[4];
    const int64_t g0 = (int64_t)blockIdx.x * PART_TILE + (int64_t)threadIdx.x * PART_ITEMS;
    uint64_t cnt = 0;
    if (g0 < w.n_bases) {
        int64_t c = first_contig(w, g0), loc;
        for (int j = 0; j < PART_ITEMS; j++) {
            const int64_t g = g0 + j;
            if (g >= w.n_bases) break;
            const uint64_t key = window_key_at(w, g, c, &loc);
            cnt += (key >= klo && key < khi) ? 1 : 0;
        }
    }
    uint64_t tot;
    (void)block_excl_scan_256(cnt, sh, tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = (uint32_t)tot;
}

// The same windows, written in position order at the block's scanned offset.
__global__ void __launch_bounds__(256) part_write_kernel(WindowScan w, uint64_t klo, uint64_t khi,
                                                         const uint64_t *__restrict__ block_off,
                                                         uint64_t *__restrict__ keys, uint64_t *__restrict__ locs) {
    __shared__ uint64_t sh[4];
    const int64_t g0 = (int64_t)blockIdx.x * PART_TILE + (int64_t)threadIdx.x * PART_ITEMS;
    uint64_t cnt = 0;
    int64_t c0 = 0;
    if (g0 < w.n_bases) {
        int64_t c = first_contig(w, g0), loc;
        c0 = c;
        for (int j = 0; j < PART_ITEMS; j++) {
            const int64_t g = g0 + j;
            if (g >= w.n_bases) break;
            const uint64_t key = window_key_at(w, g, c, &loc);
            cnt += (key >= klo && key < khi) ? 1 : 0;
        }
    }
    uint64_t tot;
    uint64_t off = block_off[blockIdx.x] + block_excl_scan_256(cnt, sh, tot);
    if (g0 < w.n_bases && cnt) {
        int64_t c = c0, loc;
        for (int j = 0; j < PART_ITEMS; j++) {
            const int64_t g = g0 + j;
            if (g >= w.n_bases) break;
            const uint64_t key = window_key_at(w, g, c, &loc);
            if (key >= klo && key < khi) {
                keys[off] = key;
                locs[off] = (uint64_t)loc;
                off++;
            }
        }
    }
}

__global__ void add_u64_kernel(uint64_t *__restrict__ dst, const uint64_t *__restrict__ src, int64_t n,
                               uint64_t add) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i] + add;
}

// Claim an empty slot of bucket `bk` for (key, val); false when it is full.
__device__ bool claim_slot(HashSlot *bk, uint64_t key, uint64_t val) {
    for (int s = 0; s < BUCKET_SLOTS; s++) {
        unsigned long long prev = atomicCAS((unsigned long long *)&bk[s].key,
                                            (unsigned long long)EMPTY_KEY, (unsigned long long)key);
        if (prev == EMPTY_KEY) {
            bk[s].val = val;
            return true;
        }
    }
    return false;
}

__device__ uint64_t slot_value(const uint64_t *offsets, int64_t i, uint64_t start_add) {
    const uint64_t start = offsets[i];
    uint64_t occ = offsets[i + 1] - start;
    if (occ > OCC_SAT) occ = OCC_SAT;
    return (occ << 40) | ((start + start_add) & START_MASK);
}

// Parallel two-choice insert of keys [0, n) (posting starts offsets[i] +
// start_add): the first bucket, then the second. A key is therefore in its
// second bucket only while its first is full (slots are never emptied, and a
// cuckoo move refills the slot it takes), which lets a lookup stop after the
// first bucket unless that bucket is full. Keys that find both full are
// listed, with their values, for the cuckoo pass.
__global__ void hash_insert_kernel(const uint64_t *__restrict__ kmers,
                                   const uint64_t *__restrict__ offsets, int64_t n_keys, uint64_t start_add,
                                   HashSlot *__restrict__ table, uint64_t n_buckets,
                                   uint64_t *__restrict__ ov_keys, uint64_t *__restrict__ ov_vals,
                                   unsigned int *__restrict__ n_overflow, unsigned int overflow_cap) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_keys;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = kmers[i];
        const uint64_t val = slot_value(offsets, i, start_add);
        const uint64_t b1 = bucket_of(key, n_buckets);
        const uint64_t b2 = bucket2_of(key, n_buckets, b1);
        if (claim_slot(table + b1 * BUCKET_SLOTS, key, val)) continue;
        if (claim_slot(table + b2 * BUCKET_SLOTS, key, val)) continue;
        const unsigned int j = atomicAdd(n_overflow, 1u);
        if (j < overflow_cap) {
            ov_keys[j] = key;
            ov_vals[j] = val;
        }
    }
}

// Cuckoo placement of the listed keys, one thread, after the parallel passes:
// a key goes into a free slot of its current bucket or displaces a resident
// (slot chosen by a step hash), which then moves to its own other bucket.
// Every loop is bounded by `step_budget` in total; running out sets *fail and
// the host rebuilds at a lower load.
__global__ void hash_cuckoo_kernel(const uint64_t *__restrict__ ov_keys, const uint64_t *__restrict__ ov_vals,
                                   HashSlot *__restrict__ table, uint64_t n_buckets,
                                   const unsigned int *__restrict__ n_overflow, unsigned int overflow_cap,
                                   int64_t step_budget, int *__restrict__ fail) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned int n = *n_overflow;
    if (n > overflow_cap) {
        *fail = 1;
        return;
    }
    int64_t steps = 0;
    for (unsigned int j = 0; j < n; j++) {
        uint64_t key = ov_keys[j];
        uint64_t val = ov_vals[j];
        uint64_t b = bucket_of(key, n_buckets);
        bool placed = false;
        while (!placed) {
            if (++steps > step_budget) {
                *fail = 1;
                return;
            }
            HashSlot *bk = table + b * BUCKET_SLOTS;
            for (int s = 0; s < BUCKET_SLOTS && !placed; s++) {
                if (bk[s].key == EMPTY_KEY) {
                    bk[s].key = key;
                    bk[s].val = val;
                    placed = true;
                }
            }
            if (placed) break;
            const int victim = (int)(splitmix_mix((uint64_t)steps ^ key) & (BUCKET_SLOTS - 1));
            const uint64_t vk = bk[victim].key, vv = bk[victim].val;
            bk[victim].key = key;
            bk[victim].val = val;
            key = vk;
            val = vv;
            const uint64_t c1 = bucket_of(key, n_buckets);
            b = (b == c1) ? bucket2_of(key, n_buckets, c1) : c1;
        }
    }
}

// Compact insert of keys [0, n) (posting starts offsets[i] + start_add):
// first choice, then second, each an 8-slot bucket claimed by a 64-bit CAS of
// the whole slot. Keys finding both full go, with their 16-byte-table value,
// to the overflow list (the stash is built from it). With `inl`, a key with
// one posting stores that posting's global base position instead of its
// start, so a lookup returns the posting itself (no dependent load): position
// coff[cid] + pos, or n_bases for a posting with cid >= n_contigs (decoded as
// an out-of-range cid, which every reader treats alike). A posting past its
// contig's end cannot be stored so and sets *inl_fail (the table is then
// rebuilt without inline postings).
__global__ void compact_insert_kernel(const uint64_t *__restrict__ kmers, const uint64_t *__restrict__ offsets,
                                      int64_t n, uint64_t start_add, uint64_t *__restrict__ ctable, int m, int r,
                                      int sb, int inl, const uint32_t *__restrict__ postings,
                                      const uint64_t *__restrict__ coff, int64_t n_contigs,
                                      uint64_t *__restrict__ ov_keys, uint64_t *__restrict__ ov_vals,
                                      unsigned int *__restrict__ n_overflow, unsigned int overflow_cap,
                                      unsigned int *__restrict__ inl_fail) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = kmers[i];
        const uint64_t start = offsets[i] + start_add;
        const uint64_t occ_all = offsets[i + 1] - offsets[i];
        const uint64_t occ = occ_all < COCC_SAT ? occ_all : COCC_SAT;
        uint64_t field = start;
        if (inl && occ_all == 1) {
            const uint32_t cid = postings[2 * start], pos = postings[2 * start + 1];
            if ((int64_t)cid >= n_contigs) {
                field = coff[n_contigs];
            } else {
                field = coff[cid] + pos;
                if (field >= coff[cid + 1]) atomicOr(inl_fail, 1u);
            }
        }
        bool done = false;
        for (int c = 0; c < 2 && !done; c++) {
            const uint64_t h = key_perm(key, m, c);
            const uint64_t b = h >> r;
            const uint64_t rem = h & ((1ull << r) - 1);
            const uint64_t slot = ((uint64_t)c << 63) | (rem << (63 - r)) | (occ << sb) | field;
            unsigned long long *bk = (unsigned long long *)(ctable + b * CBUCKET_SLOTS);
            for (int s = 0; s < CBUCKET_SLOTS && !done; s++)
                done = atomicCAS(&bk[s], 0ull, (unsigned long long)slot) == 0ull;
        }
        if (!done) {
            const unsigned int j = atomicAdd(n_overflow, 1u);
            if (j < overflow_cap) {
                ov_keys[j] = key;
                ov_vals[j] = ((occ_all < OCC_SAT ? occ_all : OCC_SAT) << 40) | (start & START_MASK);
            }
        }
    }
}

// Stash: the overflow keys in a 16-byte-slot two-choice table; keys finding
// both of their stash buckets full are listed for a cuckoo pass on the stash.
__global__ void stash_insert_kernel(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                    int64_t n, HashSlot *__restrict__ stash, uint64_t nb,
                                    uint64_t *__restrict__ ov_keys, uint64_t *__restrict__ ov_vals,
                                    unsigned int *__restrict__ n_overflow, unsigned int overflow_cap) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[i];
        const uint64_t b1 = bucket_of(key, nb);
        const uint64_t b2 = bucket2_of(key, nb, b1);
        if (claim_slot(stash + b1 * BUCKET_SLOTS, key, vals[i])) continue;
        if (claim_slot(stash + b2 * BUCKET_SLOTS, key, vals[i])) continue;
        const unsigned int j = atomicAdd(n_overflow, 1u);
        if (j < overflow_cap) {
            ov_keys[j] = key;
            ov_vals[j] = vals[i];
        }
    }
}

__global__ void fill_slots_kernel(HashSlot *t, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        t[i].key = EMPTY_KEY;
        t[i].val = 0;
    }
}

int upload(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return MA_OK;
    MA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    return MA_OK;
}

// 2-bit packing of the reference + contig offsets; fills ix->ref2/refN/coff.
int pack_reference(ma_index *ix, const uint8_t *seq, int64_t n_bases, const uint64_t *contig_off,
                   int64_t n_contigs, hipStream_t st) {
    ix->n_bases = n_bases;
    ix->n_contigs = n_contigs;
    int64_t n_words = (n_bases + 31) / 32;
    MA_TRY(ix->ref2.alloc((size_t)(n_words + REF_PAD_WORDS) * 8));
    MA_TRY(ix->refN.alloc((size_t)(n_words + REF_PAD_WORDS) * 8));
    MA_TRY(ix->coff.alloc((size_t)(n_contigs + 1) * 8));
    MA_TRY(upload(ix->coff.p, contig_off, (size_t)(n_contigs + 1) * 8, st));
    DevBuf d_seq, d_flag;
    MA_TRY(d_seq.alloc((size_t)n_bases));
    MA_TRY(d_flag.alloc(256));
    MA_TRY(upload(d_seq.p, seq, (size_t)n_bases, st));
    MA_TRY(fill_u64(st, ix->ref2.as<uint64_t>() + n_words, REF_PAD_WORDS, 0));
    MA_TRY(fill_u64(st, ix->refN.as<uint64_t>() + n_words, REF_PAD_WORDS, 0));
    MA_TRY(fill_u64(st, d_flag.as<uint64_t>(), 1, 0));
    pack_ref_kernel<<<grid_for(n_words), 256, 0, st>>>(d_seq.as<uint8_t>(), n_bases,
                                                       ix->ref2.as<uint64_t>(), ix->refN.as<uint64_t>(),
                                                       n_words, d_flag.as<int32_t>());
    MA_LAUNCHED();
    int32_t has_n = 0;
    MA_HIP(hipMemcpyAsync(&has_n, d_flag.p, 4, hipMemcpyDeviceToHost, st));
    MA_HIP(hipStreamSynchronize(st));
    ix->ref_has_n = has_n;
    return MA_OK;
}

// Free device memory once it has settled. Memory of released buffers comes
// back to the free count some milliseconds after the release returns (the
// driver frees it asynchronously): read right after the build's 200 GB of
// chunk buffers go, the count showed a third of what it reached moments later.
// So read until two readings 10 ms apart no longer grow (at most ~1 s).
int64_t free_hbm() {
    if (hipDeviceSynchronize() != hipSuccess) return 0;
    size_t fr = 0, tot = 0, prev = 0;
    for (int i = 0; i < 100; i++) {
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
        if (i > 0 && fr <= prev) break;
        prev = fr;
        usleep(10000);
    }
    return (int64_t)fr;
}

// Build decisions on stderr when MIALIGN_VERBOSE is set (table format, size,
// load, overflow counts), for sizing an index against the device's HBM.
// VLOG evaluates its arguments only then.
bool verbose() {
    static const bool on = getenv("MIALIGN_VERBOSE") != nullptr;
    return on;
}
#define VLOG(...)                          \
    do {                                   \
        if (verbose()) vlog(__VA_ARGS__); \
    } while (0)
void vlog(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[mialign index] ");
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
}

// Hash table builder. begin_* sizes and clears a table (compact 8-byte slots,
// or 16-byte slots at a load), insert adds key ranges (each with its
// posting-start shift), finish places the overflow keys (compact: the stash;
// 16-byte: the cuckoo pass) and reports whether that succeeded.
struct HashBuilder {
    static constexpr unsigned int OVERFLOW_CAP = 1u << 20;
    static constexpr int64_t STEP_BUDGET = 1ll << 22;
    static constexpr unsigned int STASH_OVERFLOW_CAP = 1u << 16;
    // Compact table: target load, and the highest load it may take (fewer
    // buckets) when the table at the target would exceed half the free HBM;
    // overflow lists are then sized for n_keys / 200 keys (8-slot buckets
    // with first-choice-first inserts overflow ~2e-3 of keys at 0.62).
    static constexpr double COMPACT_MAX_LOAD = 0.65;
    ma_index *ix = nullptr;
    hipStream_t st = nullptr;
    unsigned int ov_cap = OVERFLOW_CAP;
    DevBuf misc, ov_keys, ov_vals, ov2_keys, ov2_vals;

    int prepare(ma_index *index, hipStream_t stream, unsigned int cap = OVERFLOW_CAP) {
        ix = index;
        st = stream;
        ov_cap = cap;
        if (!misc.p) {
            MA_TRY(misc.alloc(256));
            MA_TRY(ov2_keys.alloc((size_t)STASH_OVERFLOW_CAP * 8));
            MA_TRY(ov2_vals.alloc((size_t)STASH_OVERFLOW_CAP * 8));
        }
        MA_TRY(ov_keys.reserve((size_t)ov_cap * 8));
        MA_TRY(ov_vals.reserve((size_t)ov_cap * 8));
        ix->table.release();
        ix->stash.release();
        ix->stash_nb = 0;
        return fill_u64(st, misc.as<uint64_t>(), 2, 0);  // overflow counts + fail flag
    }
    int begin_wide(ma_index *index, hipStream_t stream, int64_t n_keys, double load) {
        MA_TRY(prepare(index, stream));
        const uint64_t n_buckets = (uint64_t)((double)n_keys / (BUCKET_SLOTS * load)) + 1;
        VLOG("16-byte table: %lld keys at load %.2f (%.1f GB), free HBM %.1f GB", (long long)n_keys, load,
             (double)n_buckets * BUCKET_SLOTS * sizeof(HashSlot) / 1e9, free_hbm() / 1e9);
        ix->compact = 0;
        ix->n_buckets = n_buckets;
        ix->first_bucket_first = load <= 0.25 ? 1 : 0;
        MA_TRY(ix->table.alloc((size_t)n_buckets * BUCKET_SLOTS * sizeof(HashSlot)));
        const int64_t n_slots = (int64_t)n_buckets * BUCKET_SLOTS;
        fill_slots_kernel<<<grid_for(n_slots), 256, 0, st>>>(ix->table.as<HashSlot>(), n_slots);
        MA_LAUNCHED();
        return MA_OK;
    }
    // *fits = false when the slot fields do not fit 64 bits within a table of
    // at most half the free HBM (then the 16-byte table is used).
    // With try_inline, one-posting keys hold their position (compact_insert_kernel)
    // when the contig offsets fit the map kernel's LDS cache and the wider
    // start field needs no more buckets than the plain one.
    int begin_compact(ma_index *index, hipStream_t stream, int64_t n_keys, int64_t n_postings, int k, double load,
                      bool try_inline, bool *fits) {
        *fits = false;
        const int m = 2 * k;
        int sb = 1;
        while (sb < 40 && (1ll << sb) <= n_postings) sb++;
        int sb_pos = 1;  // global positions 0..n_bases (n_bases marks an out-of-range cid)
        while (sb_pos < 40 && (1ll << sb_pos) <= index->n_bases) sb_pos++;
        int nbits = 1;
        while (nbits < m && (double)n_keys > (double)CBUCKET_SLOTS * load * (double)(1ll << nbits)) nbits++;
        if (const char *env = getenv("MIALIGN_HASH_BUCKET_BITS")) {  // test knob: 2^bits buckets
            const int v = atoi(env);
            MA_REQUIRE(v >= 1 && v <= m && (int64_t)CBUCKET_SLOTS << v >= n_keys,
                       "MIALIGN_HASH_BUCKET_BITS out of range for this index");
            nbits = v;
        }
        while (nbits < m && (m - nbits) + 1 + COCC_BITS + sb > 64) nbits++;
        // fewer buckets (a higher load) when the table would take more than
        // half the free HBM, down to COMPACT_MAX_LOAD and the field fit
        const double budget = 0.5 * (double)free_hbm();
        unsigned int cap = OVERFLOW_CAP;
        while (nbits > 1 && (double)((int64_t)CBUCKET_SLOTS * 8 << nbits) > budget &&
               (double)n_keys <= COMPACT_MAX_LOAD * (double)CBUCKET_SLOTS * (double)(1ll << (nbits - 1)) &&
               (m - nbits + 1) + 1 + COCC_BITS + sb <= 64) {
            nbits--;
            const int64_t c = n_keys / 200;
            cap = c > (int64_t)cap ? (unsigned int)(c < (1ll << 31) ? c : (1ll << 31)) : cap;
        }
        const int r = m - nbits;
        if (r + 1 + COCC_BITS + sb > 64 || nbits > 40) return MA_OK;
        const int sb_inl = sb_pos > sb ? sb_pos : sb;
        const bool inl = try_inline && index->n_contigs < COFF_LDS && r + 1 + COCC_BITS + sb_inl <= 64 &&
                         sb_inl <= 40;
        if (inl) sb = sb_inl;
        const int64_t bytes = (int64_t)CBUCKET_SLOTS * 8 << nbits;
        VLOG("compact table: %lld keys, 2^%d buckets (%.1f GB, load %.3f), start bits %d, inline %d, "
             "overflow cap %u, half of free HBM %.1f GB%s", (long long)n_keys, nbits, bytes / 1e9,
             (double)n_keys / ((double)CBUCKET_SLOTS * (double)(1ll << nbits)), sb, (int)inl, cap, budget / 1e9,
             (double)bytes > budget ? ": does not fit" : "");
        if ((double)bytes > budget) return MA_OK;
        MA_TRY(prepare(index, stream, cap));
        ix->compact = 1;
        ix->c_m = m;
        ix->c_r = r;
        ix->c_sb = sb;
        ix->c_inline = inl ? 1 : 0;
        ix->n_buckets = 1ull << nbits;
        ix->first_bucket_first = 1;
        MA_TRY(ix->table.alloc((size_t)bytes));
        MA_TRY(fill_u64(st, ix->table.as<uint64_t>(), bytes / 8, 0));
        *fits = true;
        return MA_OK;
    }
    int insert(const uint64_t *kmers, const uint64_t *offsets, int64_t n, uint64_t start_add) {
        if (n <= 0) return MA_OK;
        if (ix->compact) {
            compact_insert_kernel<<<grid_for(n), 256, 0, st>>>(
                kmers, offsets, n, start_add, ix->table.as<uint64_t>(), ix->c_m, ix->c_r, ix->c_sb, ix->c_inline,
                ix->postings.as<uint32_t>(), ix->coff.as<uint64_t>(), ix->n_contigs, ov_keys.as<uint64_t>(),
                ov_vals.as<uint64_t>(), misc.as<unsigned int>(), ov_cap, misc.as<unsigned int>() + 3);
        } else {
            hash_insert_kernel<<<grid_for(n), 256, 0, st>>>(kmers, offsets, n, start_add, ix->table.as<HashSlot>(),
                                                            ix->n_buckets, ov_keys.as<uint64_t>(),
                                                            ov_vals.as<uint64_t>(), misc.as<unsigned int>(),
                                                            ov_cap);
        }
        MA_LAUNCHED();
        return MA_OK;
    }
    // *inl_fail: a one-posting key's posting lies past its contig (compact
    // table with inline postings only); the table must be rebuilt without.
    int finish(bool *ok, bool *inl_fail = nullptr) {
        *ok = false;
        if (inl_fail) *inl_fail = false;
        if (ix->compact) {
            unsigned int hd[4] = {0, 0, 0, 0};  // overflow count, fail flag, stash overflow, inline failure
            MA_HIP(hipMemcpyAsync(hd, misc.p, sizeof(hd), hipMemcpyDeviceToHost, st));
            MA_HIP(hipStreamSynchronize(st));
            const unsigned int n_ov = hd[0];
            ix->n_cuckoo = (int64_t)n_ov;
            VLOG("compact table: %u keys found both buckets full (cap %u), inline failure %u", n_ov, ov_cap, hd[3]);
            if (hd[3]) {
                if (inl_fail) *inl_fail = true;
                return MA_OK;
            }
            if (n_ov > ov_cap) return MA_OK;
            if (n_ov > 0) {
                const uint64_t nb = (uint64_t)((double)n_ov / (BUCKET_SLOTS * 0.25)) + 1;
                MA_TRY(ix->stash.alloc((size_t)nb * BUCKET_SLOTS * sizeof(HashSlot)));
                fill_slots_kernel<<<grid_for((int64_t)nb * BUCKET_SLOTS), 256, 0, st>>>(ix->stash.as<HashSlot>(),
                                                                                        (int64_t)nb * BUCKET_SLOTS);
                MA_LAUNCHED();
                // misc: [0] overflow count, [1] fail flag, [2] stash overflow count, [3] inline failure
                unsigned int *n_ov2 = misc.as<unsigned int>() + 2;
                stash_insert_kernel<<<grid_for(n_ov), 256, 0, st>>>(ov_keys.as<uint64_t>(), ov_vals.as<uint64_t>(),
                                                                    n_ov, ix->stash.as<HashSlot>(), nb,
                                                                    ov2_keys.as<uint64_t>(), ov2_vals.as<uint64_t>(),
                                                                    n_ov2, STASH_OVERFLOW_CAP);
                MA_LAUNCHED();
                hash_cuckoo_kernel<<<1, 64, 0, st>>>(ov2_keys.as<uint64_t>(), ov2_vals.as<uint64_t>(),
                                                     ix->stash.as<HashSlot>(), nb, n_ov2, STASH_OVERFLOW_CAP,
                                                     STEP_BUDGET, misc.as<int>() + 1);
                MA_LAUNCHED();
                ix->stash_nb = nb;
            }
        } else {
            hash_cuckoo_kernel<<<1, 64, 0, st>>>(ov_keys.as<uint64_t>(), ov_vals.as<uint64_t>(),
                                                 ix->table.as<HashSlot>(), ix->n_buckets, misc.as<unsigned int>(),
                                                 ov_cap, STEP_BUDGET, misc.as<int>() + 1);
            MA_LAUNCHED();
        }
        int h[2] = {0, 0};
        MA_HIP(hipMemcpyAsync(h, misc.p, 8, hipMemcpyDeviceToHost, st));
        MA_HIP(hipStreamSynchronize(st));
        if (!ix->compact) ix->n_cuckoo = (int64_t)h[0];
        *ok = h[1] == 0;
        return MA_OK;
    }
};

// First table load: 0.15 when that table takes at most a tenth of the free
// HBM (almost every lookup then reads one bucket), else 0.5 (lookups read both
// buckets at once); MIALIGN_HASH_LOAD in (0, 0.95] overrides it (a tuning and
// test knob: high loads send many keys through the cuckoo pass).
int first_hash_load(int64_t n_keys, double *load) {
    *load = 0.5;
    if ((double)n_keys / 0.15 * (double)sizeof(HashSlot) <= 0.1 * (double)free_hbm()) *load = 0.15;
    if (const char *env = getenv("MIALIGN_HASH_LOAD")) {
        const double v = atof(env);
        MA_REQUIRE(v > 0 && v <= 0.95, "MIALIGN_HASH_LOAD must be in (0, 0.95]");
        *load = v;
    }
    return MA_OK;
}

// The hash table, with insert_all(hb) adding every key: the compact format
// first (unless MIALIGN_HASH_FORMAT=wide, or its fields do not fit), else
// 16-byte slots at the first load, then 0.4 and 0.3. MIALIGN_HASH_LOAD sets
// the load of whichever format is built.
template <class F>
int build_table(ma_index *ix, hipStream_t st, int64_t n_keys, int64_t n_postings, F insert_all) {
    HashBuilder hb;
    const char *fmt = getenv("MIALIGN_HASH_FORMAT");
    const bool wide_only = fmt && strcmp(fmt, "wide") == 0;
    double load_env = 0;
    if (const char *env = getenv("MIALIGN_HASH_LOAD")) {
        load_env = atof(env);
        MA_REQUIRE(load_env > 0 && load_env <= 0.95, "MIALIGN_HASH_LOAD must be in (0, 0.95]");
    }
    const char *inl_env = getenv("MIALIGN_HASH_INLINE");  // test knob: "0" keeps every key's start
    bool try_inline = !(inl_env && strcmp(inl_env, "0") == 0);
    while (!wide_only) {
        bool fits = false;
        MA_TRY(hb.begin_compact(ix, st, n_keys, n_postings, ix->k, load_env > 0 ? load_env : 0.35, try_inline,
                                &fits));
        if (!fits) break;
        MA_TRY(insert_all(hb));
        bool ok = false, inl_fail = false;
        MA_TRY(hb.finish(&ok, &inl_fail));
        if (ok) return MA_OK;
        if (!inl_fail || !try_inline) break;
        try_inline = false;
    }
    double first;
    MA_TRY(first_hash_load(n_keys, &first));
    for (double load : {first, 0.4, 0.3}) {
        if (load > first) continue;
        MA_TRY(hb.begin_wide(ix, st, n_keys, load));
        MA_TRY(insert_all(hb));
        bool ok = false;
        MA_TRY(hb.finish(&ok));
        if (ok) return MA_OK;
    }
    set_error("hash index build: overflow placement did not converge");
    return MA_ERR_NOMEM;
}

// Table over the resident kmers/offsets.
int build_hash(ma_index *ix, hipStream_t st) {
    return build_table(ix, st, ix->n_keys, ix->n_postings, [&](HashBuilder &hb) {
        return hb.insert(ix->kmers.as<uint64_t>(), ix->offsets.as<uint64_t>(), ix->n_keys, 0);
    });
}

int index_alloc(int device, ma_index **out) {
    MA_REQUIRE(out, "index: null output");
    MA_HIP(hipSetDevice(device));
    ma_index *ix = new ma_index();
    ix->device = device;
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) n_cu = 0;
    ix->n_cu = n_cu > 0 ? n_cu : 256;
    hipError_t e = hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ix;
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return MA_ERR_HIP;
    }
    *out = ix;
    return MA_OK;
}

struct MappedFile {
    void *p = nullptr;
    size_t n = 0;
    ~MappedFile() {
        if (p && n) munmap(p, n);
    }
    int open(const std::string &path) {
        int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) {
            set_error("cannot open " + path);
            return MA_ERR_IO;
        }
        struct stat st;
        if (fstat(fd, &st) != 0) {
            ::close(fd);
            set_error("cannot stat " + path);
            return MA_ERR_IO;
        }
        n = (size_t)st.st_size;
        if (n > 0) {
            p = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
            if (p == MAP_FAILED) {
                p = nullptr;
                ::close(fd);
                set_error("mmap failed: " + path);
                return MA_ERR_IO;
            }
        }
        ::close(fd);
        return MA_OK;
    }
};

// Minimal meta.json reader: integer fields and the format string.
bool json_int(const std::string &s, const char *key, int64_t *out) {
    std::string pat = std::string("\"") + key + "\"";
    size_t p = s.find(pat);
    if (p == std::string::npos) return false;
    p = s.find(':', p);
    if (p == std::string::npos) return false;
    p++;
    while (p < s.size() && (s[p] == ' ' || s[p] == '\n' || s[p] == '\t')) p++;
    char *end = nullptr;
    long long v = strtoll(s.c_str() + p, &end, 10);
    if (end == s.c_str() + p) return false;
    *out = v;
    return true;
}

}  // namespace

namespace ma {

IndexView index_view(const ma_index *ix) {
    IndexView v;
    v.table = ix->table.as<HashSlot>();
    v.n_buckets = ix->n_buckets;
    v.first_bucket_first = ix->first_bucket_first;
    v.compact = ix->compact;
    v.ctable = ix->compact ? ix->table.as<uint64_t>() : nullptr;
    v.c_m = ix->c_m;
    v.c_r = ix->c_r;
    v.c_sb = ix->c_sb;
    v.c_inline = ix->compact ? ix->c_inline : 0;
    v.stash = ix->stash.as<HashSlot>();
    v.stash_nb = ix->stash_nb;
    v.k = ix->k;
    v.postings = ix->postings.as<uint32_t>();
    v.n_postings = ix->n_postings;
    v.ref2 = ix->ref2.as<uint64_t>();
    v.refN = ix->refN.as<uint64_t>();
    v.ref_has_n = ix->ref_has_n;
    v.contig_off = ix->coff.as<uint64_t>();
    v.n_contigs = ix->n_contigs;
    v.n_bases = ix->n_bases;
    return v;
}

}  // namespace ma

extern "C" {

int ma_index_from_dense(int device, int32_t k, const uint64_t *kmers, int64_t n_keys,
                        const uint64_t *offsets, const uint32_t *postings, int64_t n_postings,
                        const uint8_t *seq, int64_t n_bases, const uint64_t *contig_off,
                        int64_t n_contigs, ma_index **out) {
    MA_REQUIRE(k >= 1 && k <= 31, "ma_index_from_dense: k must be in 1..31");
    MA_REQUIRE(n_keys >= 0 && n_postings >= 0 && n_bases >= 0 && n_contigs >= 1,
               "ma_index_from_dense: bad sizes");
    MA_REQUIRE(n_postings < (1ll << 40), "ma_index_from_dense: more than 2^40 postings");
    ma_index *ix = nullptr;
    MA_TRY(index_alloc(device, &ix));
    hipStream_t st = ix->stream;
    ix->k = k;
    ix->n_keys = n_keys;
    ix->n_postings = n_postings;
    int rc = MA_OK;
    do {
        if ((rc = ix->kmers.alloc((size_t)n_keys * 8)) != MA_OK) break;
        if ((rc = ix->offsets.alloc((size_t)(n_keys + 1) * 8)) != MA_OK) break;
        if ((rc = ix->postings.alloc((size_t)n_postings * 8)) != MA_OK) break;
        if ((rc = upload(ix->kmers.p, kmers, (size_t)n_keys * 8, st)) != MA_OK) break;
        if ((rc = upload(ix->offsets.p, offsets, (size_t)(n_keys + 1) * 8, st)) != MA_OK) break;
        if ((rc = upload(ix->postings.p, postings, (size_t)n_postings * 8, st)) != MA_OK) break;
        if ((rc = pack_reference(ix, seq, n_bases, contig_off, n_contigs, st)) != MA_OK) break;
        if ((rc = build_hash(ix, st)) != MA_OK) break;
    } while (0);
    if (rc != MA_OK) {
        ma_index_close(ix);
        return rc;
    }
    *out = ix;
    return MA_OK;
}

void ma_build_opts_default(ma_build_opts *o) {
    if (!o) return;
    o->keep_dense = 1;
    o->partition_bits = 10;
    o->chunk_windows = 0;
}

int ma_index_build(int device, int32_t k, const uint8_t *seq, int64_t n_bases,
                   const uint64_t *contig_off, int64_t n_contigs, ma_index **out) {
    ma_build_opts o;
    ma_build_opts_default(&o);
    return ma_index_build_opts(device, k, seq, n_bases, contig_off, n_contigs, &o, out);
}

}  // extern "C"

namespace {

// Consecutive partitions grouped into chunks of at most `cap` windows (a
// partition larger than cap is a chunk of its own).
std::vector<std::pair<int, int>> plan_chunks(const std::vector<uint64_t> &hist, uint64_t cap) {
    std::vector<std::pair<int, int>> chunks;
    int b = 0;
    const int nb = (int)hist.size();
    while (b < nb) {
        uint64_t sum = hist[b];
        int e = b + 1;
        while (e < nb && sum + hist[e] <= cap) sum += hist[e++];
        chunks.push_back({b, e});
        b = e;
    }
    return chunks;
}

// Resident buffers of one chunk: collected pairs, sort alternates, radix and
// scan scratch, and the chunk's CSR (kmers / offsets relative to the chunk).
struct ChunkBufs {
    DevBuf K0, K1, V0, V1, heads, excl, part, bh, bo, h8, km, of;
    int reserve(int64_t n) {
        const int64_t m = n > 0 ? n : 1;
        const int64_t nb = rs_nblocks(m);
        MA_TRY(K0.reserve((size_t)m * 8));
        MA_TRY(K1.reserve((size_t)m * 8));
        MA_TRY(V0.reserve((size_t)m * 8));
        MA_TRY(V1.reserve((size_t)m * 8));
        MA_TRY(heads.reserve((size_t)m * 4));
        MA_TRY(excl.reserve((size_t)m * 8));
        MA_TRY(part.reserve(scan_scratch_bytes(m)));
        MA_TRY(bh.reserve((size_t)(256 * nb * 4)));
        MA_TRY(bo.reserve((size_t)(256 * nb * 8)));
        MA_TRY(h8.reserve(8 * 256 * 8));
        MA_TRY(km.reserve((size_t)m * 8));
        MA_TRY(of.reserve((size_t)(m + 1) * 8));
        return MA_OK;
    }
    void release() {
        for (DevBuf *b : {&K0, &K1, &V0, &V1, &heads, &excl, &part, &bh, &bo, &h8, &km, &of}) b->release();
    }
};
constexpr int64_t CHUNK_BYTES_PER_WINDOW = 8 * 4 + 4 + 8 + 8 + 16 + 1;

// Collect the windows of key range [klo, khi) in position order and sort
// them by key (stable). Returns the window count through *nc; sorted keys
// and locations end in (*kp, *vp).
int collect_sorted(const WindowScan &w, uint64_t klo, uint64_t khi, int64_t nc, ChunkBufs &cb, DevBuf &blk,
                   DevBuf &blk_off, DevBuf &blk_part, int64_t nblk, hipStream_t st, uint64_t **kp, uint64_t **vp) {
    part_count_kernel<<<(unsigned)nblk, 256, 0, st>>>(w, klo, khi, blk.as<uint32_t>());
    MA_LAUNCHED();
    MA_TRY(exclusive_scan_u32(st, nblk, blk.as<uint32_t>(), blk_off.as<uint64_t>(), blk_part.as<uint64_t>(),
                              nullptr));
    part_write_kernel<<<(unsigned)nblk, 256, 0, st>>>(w, klo, khi, blk_off.as<uint64_t>(), cb.K0.as<uint64_t>(),
                                                      cb.V0.as<uint64_t>());
    MA_LAUNCHED();
    uint64_t *k0 = cb.K0.as<uint64_t>(), *k1 = cb.K1.as<uint64_t>();
    uint64_t *v0 = cb.V0.as<uint64_t>(), *v1 = cb.V1.as<uint64_t>();
    RadixScratch rs{cb.bh.as<uint32_t>(), cb.bo.as<uint64_t>(), cb.h8.as<unsigned long long>()};
    MA_TRY(radix_sort_pairs<uint64_t>(st, nc, k0, v0, k1, v1, rs, 2 * w.k));
    *kp = k0;
    *vp = v0;
    return MA_OK;
}

// Heads + exclusive scan over nc sorted keys; returns the unique-key count.
int count_keys(int64_t nc, const uint64_t *kp, ChunkBufs &cb, hipStream_t st, int64_t *n_keys) {
    *n_keys = 0;
    if (nc <= 0) return MA_OK;
    key_heads_kernel<<<grid_for(nc), 256, 0, st>>>(kp, nc, cb.heads.as<uint32_t>());
    MA_LAUNCHED();
    uint64_t *partial = cb.part.as<uint64_t>();
    MA_TRY(exclusive_scan_u32(st, nc, cb.heads.as<uint32_t>(), cb.excl.as<uint64_t>(), partial,
                              partial + rs_nblocks(nc)));
    uint64_t nk = 0;
    MA_HIP(hipMemcpyAsync(&nk, partial + rs_nblocks(nc), 8, hipMemcpyDeviceToHost, st));
    MA_HIP(hipStreamSynchronize(st));
    *n_keys = (int64_t)nk;
    return MA_OK;
}

// The dense-v1 arrays and the hash table of the packed reference in ix, built
// one key-range chunk at a time. Sweep 1 counts the unique keys (to size the
// table); sweep 2 writes each chunk's postings slice and inserts its keys.
int build_chunked(ma_index *ix, int k, const ma_build_opts &o, hipStream_t st) {
    const WindowScan w{ix->ref2.as<uint64_t>(), ix->refN.as<uint64_t>(), ix->coff.as<uint64_t>(), ix->n_contigs,
                       ix->n_bases, k};
    const int bits = o.partition_bits < 2 * k ? o.partition_bits : 2 * k;
    const int shift = 2 * k - bits;
    const int nbins = 1 << bits;
    const int64_t nblk = (ix->n_bases + PART_TILE - 1) / PART_TILE;
    MA_REQUIRE(nblk < (1ll << 31), "index build: reference too long");
    DevBuf d_hist, blk, blk_off, blk_part;
    MA_TRY(d_hist.alloc((size_t)nbins * 8));
    MA_TRY(blk.alloc((size_t)nblk * 4));
    MA_TRY(blk_off.alloc((size_t)nblk * 8));
    MA_TRY(blk_part.alloc(scan_scratch_bytes(nblk)));
    MA_TRY(fill_u64(st, d_hist.as<uint64_t>(), nbins, 0));
    part_hist_kernel<<<(unsigned)nblk, 256, 0, st>>>(w, shift, nbins, d_hist.as<unsigned long long>());
    MA_LAUNCHED();
    std::vector<uint64_t> hist((size_t)nbins);
    MA_HIP(hipMemcpyAsync(hist.data(), d_hist.p, (size_t)nbins * 8, hipMemcpyDeviceToHost, st));
    MA_HIP(hipStreamSynchronize(st));
    uint64_t n_valid = 0, max_bin = 0;
    for (uint64_t h : hist) {
        n_valid += h;
        max_bin = h > max_bin ? h : max_bin;
    }
    ix->n_postings = (int64_t)n_valid;
    auto cap_for = [&](int64_t reserve_bytes) -> uint64_t {
        if (o.chunk_windows > 0) return (uint64_t)o.chunk_windows;
        int64_t room = (int64_t)(0.85 * (double)free_hbm()) - reserve_bytes;
        int64_t c = room / CHUNK_BYTES_PER_WINDOW;
        return c > 0 ? (uint64_t)c : 0;
    };
    auto key_lo = [&](int b) -> uint64_t { return (uint64_t)b << shift; };

    VLOG("chunked build: %lld valid windows, largest partition %llu, free HBM %.1f GB", (long long)n_valid,
         (unsigned long long)max_bin, free_hbm() / 1e9);
    // sweep 1: unique keys per chunk
    ChunkBufs cb;
    int64_t n_keys = 0;
    {
        uint64_t cap = cap_for(0);
        if (cap < max_bin) cap = max_bin;  // a partition never splits (fails below if it cannot fit)
        const auto chunks = plan_chunks(hist, cap);
        for (const auto &ch : chunks) {
            uint64_t nc = 0;
            for (int b = ch.first; b < ch.second; b++) nc += hist[b];
            if (nc == 0) continue;
            MA_TRY(cb.reserve((int64_t)nc));
            uint64_t *kp, *vp;
            MA_TRY(collect_sorted(w, key_lo(ch.first), key_lo(ch.second), (int64_t)nc, cb, blk, blk_off, blk_part,
                                  nblk, st, &kp, &vp));
            int64_t nk = 0;
            MA_TRY(count_keys((int64_t)nc, kp, cb, st, &nk));
            n_keys += nk;
        }
    }
    ix->n_keys = n_keys;
    VLOG("sweep 1: %lld keys, free HBM %.1f GB with its chunk buffers", (long long)n_keys, free_hbm() / 1e9);
    cb.release();  // sweep-1 buffers go before the resident arrays are allocated
    MA_TRY(ix->postings.alloc((size_t)(n_valid > 0 ? n_valid : 1) * 8));
    VLOG("postings allocated (%.1f GB), free HBM %.1f GB", (double)n_valid * 8 / 1e9, free_hbm() / 1e9);
    if (o.keep_dense) {
        MA_TRY(ix->kmers.alloc((size_t)(n_keys > 0 ? n_keys : 1) * 8));
        MA_TRY(ix->offsets.alloc((size_t)(n_keys + 1) * 8));
    }
    // sweep 2 (per table attempt): postings slices, table inserts, dense-v1 slices
    return build_table(ix, st, n_keys, (int64_t)n_valid, [&](HashBuilder &hb) -> int {
        uint64_t cap = cap_for(0);
        if (cap < max_bin) cap = max_bin;
        const auto chunks = plan_chunks(hist, cap);
        uint64_t post_base = 0, key_base = 0;
        for (const auto &ch : chunks) {
            uint64_t nc = 0;
            for (int b = ch.first; b < ch.second; b++) nc += hist[b];
            if (nc == 0) continue;
            MA_TRY(cb.reserve((int64_t)nc));
            uint64_t *kp, *vp;
            MA_TRY(collect_sorted(w, key_lo(ch.first), key_lo(ch.second), (int64_t)nc, cb, blk, blk_off, blk_part,
                                  nblk, st, &kp, &vp));
            int64_t nk = 0;
            MA_TRY(count_keys((int64_t)nc, kp, cb, st, &nk));
            csr_kernel<<<grid_for((int64_t)nc), 256, 0, st>>>(kp, vp, cb.heads.as<uint32_t>(), cb.excl.as<uint64_t>(),
                                                              (int64_t)nc, cb.km.as<uint64_t>(), cb.of.as<uint64_t>(),
                                                              ix->postings.as<uint32_t>() + 2 * post_base);
            MA_LAUNCHED();
            set_u64_kernel<<<1, 1, 0, st>>>(cb.of.as<uint64_t>() + nk, nc);
            MA_LAUNCHED();
            MA_TRY(hb.insert(cb.km.as<uint64_t>(), cb.of.as<uint64_t>(), nk, post_base));
            if (o.keep_dense) {
                MA_HIP(hipMemcpyAsync(ix->kmers.as<uint64_t>() + key_base, cb.km.p, (size_t)nk * 8,
                                      hipMemcpyDeviceToDevice, st));
                add_u64_kernel<<<grid_for(nk), 256, 0, st>>>(ix->offsets.as<uint64_t>() + key_base,
                                                             cb.of.as<uint64_t>(), nk, post_base);
                MA_LAUNCHED();
            }
            post_base += nc;
            key_base += (uint64_t)nk;
        }
        if (o.keep_dense) {
            set_u64_kernel<<<1, 1, 0, st>>>(ix->offsets.as<uint64_t>() + n_keys, n_valid);
            MA_LAUNCHED();
        }
        return MA_OK;
    });
}

}  // namespace

extern "C" {

int ma_index_build_opts(int device, int32_t k, const uint8_t *seq, int64_t n_bases, const uint64_t *contig_off,
                        int64_t n_contigs, const ma_build_opts *opts, ma_index **out) {
    MA_REQUIRE(k >= 1 && k <= 31, "ma_index_build: k must be in 1..31");
    MA_REQUIRE(n_bases > 0 && n_contigs >= 1 && seq && contig_off && out, "ma_index_build: bad arguments");
    MA_REQUIRE((int64_t)contig_off[n_contigs] == n_bases, "ma_index_build: contig_off[n] != n_bases");
    MA_REQUIRE(n_contigs < (1ll << 32), "ma_index_build: too many contigs");
    ma_build_opts o;
    ma_build_opts_default(&o);
    if (opts) o = *opts;
    MA_REQUIRE(o.partition_bits >= 1 && o.partition_bits <= 10 && o.chunk_windows >= 0,
               "ma_index_build: partition_bits must be in 1..10, chunk_windows >= 0");
    ma_index *ix = nullptr;
    MA_TRY(index_alloc(device, &ix));
    hipStream_t st = ix->stream;
    ix->k = k;
    ix->keep_dense = o.keep_dense ? 1 : 0;
    int rc = pack_reference(ix, seq, n_bases, contig_off, n_contigs, st);
    if (rc == MA_OK) rc = build_chunked(ix, k, o, st);
    if (rc == MA_OK && hipStreamSynchronize(st) != hipSuccess) {
        set_error("index build failed on device");
        rc = MA_ERR_HIP;
    }
    if (rc != MA_OK) {
        ma_index_close(ix);
        return rc;
    }
    *out = ix;
    return MA_OK;
}

int ma_index_open(const char *dir, int device, ma_index **out) {
    MA_REQUIRE(dir && out, "ma_index_open: bad arguments");
    std::string d(dir);
    MappedFile meta, km, of, po, sq, co;
    MA_TRY(meta.open(d + "/meta.json"));
    std::string ms((const char *)meta.p, meta.n);
    if (ms.find("dense-v1") == std::string::npos) {
        set_error("ma_index_open: meta.json is not format dense-v1: " + d);
        return MA_ERR_IO;
    }
    int64_t k = 0, n_keys = 0, n_post = 0;
    if (!json_int(ms, "k", &k) || !json_int(ms, "n_keys", &n_keys) ||
        !json_int(ms, "n_postings", &n_post)) {
        set_error("ma_index_open: meta.json lacks k / n_keys / n_postings");
        return MA_ERR_IO;
    }
    MA_TRY(km.open(d + "/kmers.bin"));
    MA_TRY(of.open(d + "/offsets.bin"));
    MA_TRY(po.open(d + "/postings.bin"));
    MA_TRY(sq.open(d + "/sequences.bin"));
    MA_TRY(co.open(d + "/contig_offsets.bin"));
    int64_t n_contigs = (int64_t)co.n / 8 - 1;
    // Size validation as in linear_gpu_locate.mojo:1621-1628.
    if ((int64_t)km.n < n_keys * 8 || (int64_t)of.n < (n_keys + 1) * 8 || (int64_t)po.n < n_post * 8 ||
        n_contigs < 1) {
        set_error("ma_index_open: mmap smaller than expected tables");
        return MA_ERR_IO;
    }
    const uint64_t *coff = (const uint64_t *)co.p;
    if ((int64_t)coff[n_contigs] > (int64_t)sq.n) {
        set_error("ma_index_open: sequences.bin shorter than contig_offsets");
        return MA_ERR_IO;
    }
    return ma_index_from_dense(device, (int32_t)k, (const uint64_t *)km.p, n_keys,
                               (const uint64_t *)of.p, (const uint32_t *)po.p, n_post,
                               (const uint8_t *)sq.p, (int64_t)coff[n_contigs], coff, n_contigs, out);
}

int ma_index_stats(const ma_index *ix, int64_t *n_keys, int64_t *n_postings, int64_t *n_bases,
                   int64_t *n_contigs, int64_t *device_bytes, int32_t *k) {
    MA_REQUIRE(ix, "ma_index_stats: null index");
    if (n_keys) *n_keys = ix->n_keys;
    if (n_postings) *n_postings = ix->n_postings;
    if (n_bases) *n_bases = ix->n_bases;
    if (n_contigs) *n_contigs = ix->n_contigs;
    if (k) *k = ix->k;
    if (device_bytes)
        *device_bytes = (int64_t)(ix->kmers.bytes + ix->offsets.bytes + ix->postings.bytes + ix->ref2.bytes +
                                  ix->refN.bytes + ix->coff.bytes + ix->table.bytes + ix->stash.bytes);
    return MA_OK;
}

int ma_index_hash_stats(const ma_index *ix, int64_t *n_buckets, int64_t *n_overflow, int32_t *slot_bytes,
                        int32_t *inline_postings) {
    MA_REQUIRE(ix, "ma_index_hash_stats: null index");
    if (n_buckets) *n_buckets = (int64_t)ix->n_buckets;
    if (n_overflow) *n_overflow = ix->n_cuckoo;
    if (slot_bytes) *slot_bytes = ix->compact ? 8 : (int32_t)sizeof(HashSlot);
    if (inline_postings) *inline_postings = ix->compact ? ix->c_inline : 0;
    return MA_OK;
}

int ma_index_export_dense(ma_index *ix, uint64_t *kmers, uint64_t *offsets, uint32_t *postings) {
    MA_REQUIRE(ix, "ma_index_export_dense: null index");
    if (!ix->keep_dense && (kmers || offsets)) {
        set_error("ma_index_export_dense: index built without resident kmers/offsets (keep_dense = 0)");
        return MA_ERR_UNSUPPORTED;
    }
    MA_HIP(hipSetDevice(ix->device));
    hipStream_t st = ix->stream;
    if (kmers && ix->n_keys)
        MA_HIP(hipMemcpyAsync(kmers, ix->kmers.p, (size_t)ix->n_keys * 8, hipMemcpyDeviceToHost, st));
    if (offsets)
        MA_HIP(hipMemcpyAsync(offsets, ix->offsets.p, (size_t)(ix->n_keys + 1) * 8, hipMemcpyDeviceToHost, st));
    if (postings && ix->n_postings)
        MA_HIP(hipMemcpyAsync(postings, ix->postings.p, (size_t)ix->n_postings * 8, hipMemcpyDeviceToHost, st));
    MA_HIP(hipStreamSynchronize(st));
    return MA_OK;
}

int ma_index_close(ma_index *ix) {
    if (!ix) return MA_OK;
    (void)hipSetDevice(ix->device);
    if (ix->stream) (void)hipStreamDestroy(ix->stream);
    delete ix;
    return MA_OK;
}

}  // extern "C"
