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

// One key per base position; windows that leave the contig or hold an N get
// the sentinel key 1 << 2k (sorts after every valid key).
__global__ void window_keys_kernel(const uint64_t *__restrict__ ref2, const uint64_t *__restrict__ refN,
                                   const uint64_t *__restrict__ contig_off, int64_t n_contigs,
                                   int64_t n_bases, int k, uint64_t *__restrict__ keys,
                                   uint64_t *__restrict__ locs, unsigned long long *__restrict__ n_valid) {
    uint64_t invalid = 1ull << (2 * k);
    unsigned long long local = 0;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n_bases;
         g += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = n_contigs;  // last contig whose start <= g
        while (hi - lo > 1) {
            int64_t mid = (lo + hi) >> 1;
            if ((int64_t)contig_off[mid] <= g) lo = mid;
            else hi = mid;
        }
        int64_t p = g - (int64_t)contig_off[lo];
        int64_t clen = (int64_t)(contig_off[lo + 1] - contig_off[lo]);
        uint64_t key = invalid;
        if (p + k <= clen) {
            uint64_t nw = window32(refN, g) >> (64 - 2 * k);
            if (nw == 0) {
                key = window32(ref2, g) >> (64 - 2 * k);
                local++;
            }
        }
        keys[g] = key;
        locs[g] = ((uint64_t)lo << 32) | (uint64_t)p;
    }
    if (local) atomicAdd(n_valid, local);
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

__device__ uint64_t slot_value(const uint64_t *offsets, int64_t i) {
    uint64_t start = offsets[i];
    uint64_t occ = offsets[i + 1] - start;
    if (occ > OCC_SAT) occ = OCC_SAT;
    return (occ << 40) | (start & START_MASK);
}

// Parallel two-choice insert: the less occupied of a key's two buckets first
// (a racy hint, exactness does not depend on it), then the other. Keys that
// find both full are listed for the cuckoo pass.
__global__ void hash_insert_kernel(const uint64_t *__restrict__ kmers,
                                   const uint64_t *__restrict__ offsets, int64_t n_keys,
                                   HashSlot *__restrict__ table, uint64_t n_buckets,
                                   int64_t *__restrict__ overflow, unsigned int *__restrict__ n_overflow,
                                   unsigned int overflow_cap) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_keys;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = kmers[i];
        const uint64_t val = slot_value(offsets, i);
        const uint64_t b1 = bucket_of(key, n_buckets);
        const uint64_t b2 = bucket2_of(key, n_buckets, b1);
        int o1 = 0, o2 = 0;
        for (int s = 0; s < BUCKET_SLOTS; s++) {
            o1 += table[b1 * BUCKET_SLOTS + s].key != EMPTY_KEY;
            o2 += table[b2 * BUCKET_SLOTS + s].key != EMPTY_KEY;
        }
        const uint64_t first = o2 < o1 ? b2 : b1, second = o2 < o1 ? b1 : b2;
        if (claim_slot(table + first * BUCKET_SLOTS, key, val)) continue;
        if (claim_slot(table + second * BUCKET_SLOTS, key, val)) continue;
        unsigned int j = atomicAdd(n_overflow, 1u);
        if (j < overflow_cap) overflow[j] = i;
    }
}

// Cuckoo placement of the listed keys, one thread, after the parallel pass:
// a key goes into a free slot of its current bucket or displaces a resident
// (slot chosen by a step hash), which then moves to its own other bucket.
// Every loop is bounded by `step_budget` in total; running out sets *fail and
// the host rebuilds with more buckets.
__global__ void hash_cuckoo_kernel(const uint64_t *__restrict__ kmers, const uint64_t *__restrict__ offsets,
                                   HashSlot *__restrict__ table, uint64_t n_buckets,
                                   const int64_t *__restrict__ overflow, const unsigned int *__restrict__ n_overflow,
                                   unsigned int overflow_cap, int64_t step_budget, int *__restrict__ fail) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned int n = *n_overflow;
    if (n > overflow_cap) {
        *fail = 1;
        return;
    }
    int64_t steps = 0;
    for (unsigned int j = 0; j < n; j++) {
        uint64_t key = kmers[overflow[j]];
        uint64_t val = slot_value(offsets, overflow[j]);
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

// Two-choice hash table over the resident kmers/offsets at load 0.5 (0.4,
// then 0.3, if the cuckoo pass runs out of steps). MIALIGN_HASH_LOAD in
// (0, 0.95] sets the first load (a tuning and test knob: high loads send many
// keys through the cuckoo pass).
int build_hash(ma_index *ix, hipStream_t st) {
    const int64_t nk = ix->n_keys;
    constexpr unsigned int OVERFLOW_CAP = 1u << 16;
    constexpr int64_t STEP_BUDGET = 1ll << 22;
    DevBuf d_misc, d_over;
    MA_TRY(d_misc.alloc(256));
    MA_TRY(d_over.alloc((size_t)OVERFLOW_CAP * 8));
    unsigned int *d_nover = d_misc.as<unsigned int>();
    int *d_fail = d_misc.as<int>() + 1;
    double first = 0.5;
    if (const char *env = getenv("MIALIGN_HASH_LOAD")) {
        const double v = atof(env);
        MA_REQUIRE(v > 0 && v <= 0.95, "MIALIGN_HASH_LOAD must be in (0, 0.95]");
        first = v;
    }
    for (double load : {first, 0.4, 0.3}) {
        if (load > first) continue;
        const uint64_t n_buckets = (uint64_t)((double)nk / (BUCKET_SLOTS * load)) + 1;
        ix->n_buckets = n_buckets;
        MA_TRY(ix->table.alloc((size_t)n_buckets * BUCKET_SLOTS * sizeof(HashSlot)));
        const int64_t n_slots = (int64_t)n_buckets * BUCKET_SLOTS;
        fill_slots_kernel<<<grid_for(n_slots), 256, 0, st>>>(ix->table.as<HashSlot>(), n_slots);
        MA_LAUNCHED();
        MA_TRY(fill_u64(st, d_misc.as<uint64_t>(), 1, 0));
        if (nk > 0) {
            hash_insert_kernel<<<grid_for(nk), 256, 0, st>>>(ix->kmers.as<uint64_t>(), ix->offsets.as<uint64_t>(),
                                                             nk, ix->table.as<HashSlot>(), n_buckets,
                                                             d_over.as<int64_t>(), d_nover, OVERFLOW_CAP);
            MA_LAUNCHED();
            hash_cuckoo_kernel<<<1, 64, 0, st>>>(ix->kmers.as<uint64_t>(), ix->offsets.as<uint64_t>(),
                                                 ix->table.as<HashSlot>(), n_buckets, d_over.as<int64_t>(),
                                                 d_nover, OVERFLOW_CAP, STEP_BUDGET, d_fail);
            MA_LAUNCHED();
        }
        int h[2] = {0, 0};
        MA_HIP(hipMemcpyAsync(h, d_misc.p, 8, hipMemcpyDeviceToHost, st));
        MA_HIP(hipStreamSynchronize(st));
        ix->n_cuckoo = (int64_t)h[0];
        if (!h[1]) return MA_OK;
    }
    set_error("hash index build: cuckoo placement did not converge");
    return MA_ERR_NOMEM;
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

int ma_index_build(int device, int32_t k, const uint8_t *seq, int64_t n_bases,
                   const uint64_t *contig_off, int64_t n_contigs, ma_index **out) {
    MA_REQUIRE(k >= 1 && k <= 31, "ma_index_build: k must be in 1..31");
    MA_REQUIRE(n_bases > 0 && n_contigs >= 1 && seq && contig_off, "ma_index_build: bad arguments");
    MA_REQUIRE((int64_t)contig_off[n_contigs] == n_bases, "ma_index_build: contig_off[n] != n_bases");
    MA_REQUIRE(n_contigs < (1ll << 32), "ma_index_build: too many contigs");
    ma_index *ix = nullptr;
    MA_TRY(index_alloc(device, &ix));
    hipStream_t st = ix->stream;
    ix->k = k;
    int rc = MA_OK;
    do {
        if ((rc = pack_reference(ix, seq, n_bases, contig_off, n_contigs, st)) != MA_OK) break;
        int64_t n = n_bases;
        DevBuf K0, K1, V0, V1, heads, excl, partial, bh, bo, h8, cnt;
        int64_t nb = rs_nblocks(n);
        if ((rc = K0.alloc((size_t)n * 8)) || (rc = K1.alloc((size_t)n * 8)) ||
            (rc = V0.alloc((size_t)n * 8)) || (rc = V1.alloc((size_t)n * 8)) ||
            (rc = bh.alloc((size_t)(256 * nb * 4))) || (rc = bo.alloc((size_t)(256 * nb * 8))) ||
            (rc = h8.alloc(8 * 256 * 8)) || (rc = cnt.alloc(256)))
            break;
        if ((rc = fill_u64(st, cnt.as<uint64_t>(), 1, 0)) != MA_OK) break;
        window_keys_kernel<<<grid_for(n), 256, 0, st>>>(
            ix->ref2.as<uint64_t>(), ix->refN.as<uint64_t>(), ix->coff.as<uint64_t>(), n_contigs,
            n_bases, k, K0.as<uint64_t>(), V0.as<uint64_t>(), cnt.as<unsigned long long>());
        if (hipGetLastError() != hipSuccess) {
            set_error("window_keys_kernel launch failed");
            rc = MA_ERR_HIP;
            break;
        }
        unsigned long long n_valid = 0;
        if (hipMemcpyAsync(&n_valid, cnt.p, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("index build: count readback failed");
            rc = MA_ERR_HIP;
            break;
        }
        RadixScratch rs{bh.as<uint32_t>(), bo.as<uint64_t>(), h8.as<unsigned long long>()};
        uint64_t *kp = K0.as<uint64_t>(), *ka = K1.as<uint64_t>();
        uint64_t *vp = V0.as<uint64_t>(), *va = V1.as<uint64_t>();
        if ((rc = radix_sort_pairs<uint64_t>(st, n, kp, vp, ka, va, rs, 2 * k + 1)) != MA_OK) break;
        bh.release();
        bo.release();
        int64_t nv = (int64_t)n_valid;
        ix->n_postings = nv;
        if ((rc = heads.alloc((size_t)(nv > 0 ? nv : 1) * 4)) || (rc = excl.alloc((size_t)(nv > 0 ? nv : 1) * 8)) ||
            (rc = partial.alloc(scan_scratch_bytes(nv > 0 ? nv : 1))))
            break;
        uint64_t n_keys = 0;
        if (nv > 0) {
            key_heads_kernel<<<grid_for(nv), 256, 0, st>>>(kp, nv, heads.as<uint32_t>());
            if ((rc = exclusive_scan_u32(st, nv, heads.as<uint32_t>(), excl.as<uint64_t>(),
                                         partial.as<uint64_t>(), partial.as<uint64_t>() + rs_nblocks(nv))) != MA_OK)
                break;
            if (hipMemcpyAsync(&n_keys, partial.as<uint64_t>() + rs_nblocks(nv), 8,
                               hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                set_error("index build: key count readback failed");
                rc = MA_ERR_HIP;
                break;
            }
        }
        ix->n_keys = (int64_t)n_keys;
        if ((rc = ix->kmers.alloc((size_t)n_keys * 8)) || (rc = ix->offsets.alloc((size_t)(n_keys + 1) * 8)) ||
            (rc = ix->postings.alloc((size_t)nv * 8)))
            break;
        if (nv > 0) {
            csr_kernel<<<grid_for(nv), 256, 0, st>>>(kp, vp, heads.as<uint32_t>(), excl.as<uint64_t>(), nv,
                                                     ix->kmers.as<uint64_t>(), ix->offsets.as<uint64_t>(),
                                                     ix->postings.as<uint32_t>());
        }
        set_u64_kernel<<<1, 1, 0, st>>>(ix->offsets.as<uint64_t>() + n_keys, (uint64_t)nv);
        if (hipGetLastError() != hipSuccess) {
            set_error("csr_kernel launch failed");
            rc = MA_ERR_HIP;
            break;
        }
        if (hipStreamSynchronize(st) != hipSuccess) {
            set_error("index build failed on device");
            rc = MA_ERR_HIP;
            break;
        }
    } while (0);
    if (rc == MA_OK) rc = build_hash(ix, st);
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
        *device_bytes = (int64_t)(ix->kmers.bytes + ix->offsets.bytes + ix->postings.bytes +
                                  ix->ref2.bytes + ix->refN.bytes + ix->coff.bytes + ix->table.bytes);
    return MA_OK;
}

int ma_index_hash_stats(const ma_index *ix, int64_t *n_buckets, int64_t *n_cuckoo) {
    MA_REQUIRE(ix, "ma_index_hash_stats: null index");
    if (n_buckets) *n_buckets = (int64_t)ix->n_buckets;
    if (n_cuckoo) *n_cuckoo = ix->n_cuckoo;
    return MA_OK;
}

int ma_index_export_dense(ma_index *ix, uint64_t *kmers, uint64_t *offsets, uint32_t *postings) {
    MA_REQUIRE(ix, "ma_index_export_dense: null index");
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
