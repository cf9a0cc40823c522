// Shared host/device helpers for the mialign library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>

#include "../../include/mialign.h"

namespace ma {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
// Current value of a launch knob (ma_tune_set / ma_tune_get).
int64_t tune_get(int knob);
// Byte offsets, inside ma_sort_markdup_dev's scratch for n records, of the two
// u32 words that are nonzero when one of its sorts gave up waiting on a tile.
void sort_error_words(int64_t n, size_t *dup_word, size_t *coord_word);

#define MA_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            ::ma::set_error(std::string(#call) + ": " + hipGetErrorString(_e) + " at " + \
                            __FILE__ + ":" + std::to_string(__LINE__));                \
            return MA_ERR_HIP;                                                         \
        }                                                                              \
    } while (0)

#define MA_REQUIRE(cond, msg)            \
    do {                                 \
        if (!(cond)) {                   \
            ::ma::set_error(msg);        \
            return MA_ERR_ARG;           \
        }                                \
    } while (0)

#define MA_TRY(call)           \
    do {                       \
        int _rc = (call);      \
        if (_rc != MA_OK) return _rc; \
    } while (0)

// Launch check: reports the launch error of the kernel just enqueued.
#define MA_LAUNCHED() MA_HIP(hipGetLastError())

// ---------------------------------------------------------- device memory
// Owning device allocation (freed on destruction).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t n) {
        release();
        if (n == 0) n = 256;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("hipMalloc(") + std::to_string(n) + "): " + hipGetErrorString(e));
            return MA_ERR_NOMEM;
        }
        bytes = n;
        return MA_OK;
    }
    // Grow-only reuse.
    int reserve(size_t n) { return (n <= bytes && p) ? MA_OK : alloc(n); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// ------------------------------------------------------------- constants
constexpr int WAVE = 64;
constexpr uint64_t LOW2 = 0x5555555555555555ull;  // low bit of each 2-bit group
constexpr uint64_t EMPTY_KEY = ~0ull;

// Hash index slot: key (MSB-first 2-bit k-mer) and value = occ(24) << 40 | start(40).
// Two-choice bucketed table: every key sits in one of its two 64-byte buckets
// (bucket_of / bucket2_of), and in the second only while the first is full.
// At low load a lookup reads the first bucket and, rarely, the second; at
// high load it loads both at once (one memory round trip either way).
struct HashSlot {
    uint64_t key;
    uint64_t val;
};
constexpr int BUCKET_SLOTS = 4;  // 64-byte buckets
// Zero words after the packed reference: a window's words are loaded as one
// batch of up to 9 words from its first word, wherever it ends.
constexpr int64_t REF_PAD_WORDS = 16;
constexpr uint64_t START_MASK = (1ull << 40) - 1;
constexpr uint64_t OCC_SAT = (1ull << 24) - 1;

// Compact table: 64-byte buckets of eight 8-byte slots. A key's slot is
// tag | occ << sb | start, the tag being the choice bit and the low r bits of
// the permuted key (the bucket index holds the rest), so equal (bucket, tag)
// means equal key. Empty slots are 0 (a key has occ >= 1).
constexpr int CBUCKET_SLOTS = 8;
constexpr int COCC_BITS = 18;  // saturates above any max_occ (< 2^17)
constexpr uint64_t COCC_SAT = (1ull << COCC_BITS) - 1;
// Contig offsets fit the map kernel's LDS cache below this contig count; the
// compact table holds one-posting keys inline only then (their decode is a
// binary search over those offsets).
constexpr int COFF_LDS = 256;

// A bijection on m-bit keys (1 <= m <= 62): two multiply (odd constant) +
// xorshift rounds modulo 2^m; choice 0 and 1 use different multipliers.
__host__ __device__ inline uint64_t key_perm(uint64_t x, int m, int c) {
    const uint64_t M = (1ull << m) - 1;
    const uint64_t c1 = c ? 0xD6E8FEB86659FD93ull : 0x9E3779B97F4A7C15ull;
    const uint64_t c2 = c ? 0xA0761D6478BD642Full : 0xBF58476D1CE4E5B9ull;
    const int s1 = (m + 1) / 2, s2 = m / 3 + 1;
    x = (x * c1) & M;
    x ^= x >> s1;
    x = (x * c2) & M;
    x ^= x >> s2;
    return x;
}

// Resident index (device pointers) — passed by value to kernels.
struct IndexView {
    const HashSlot *table;   // 16-byte-slot table (compact == 0)
    uint64_t n_buckets;
    int32_t first_bucket_first;  // 1: second bucket only when the first is full and misses
    int32_t compact;             // 1: 8-byte-slot table in ctable (+ stash)
    const uint64_t *ctable;
    int32_t c_m, c_r, c_sb;      // key bits, remainder bits, start bits
    int32_t c_inline;            // 1: a compact slot with occ 1 holds its posting's global base
                                 // position contig_off[cid] + pos (n_bases for a cid >= n_contigs)
    const HashSlot *stash;       // keys whose two compact buckets were full (16-byte slots)
    uint64_t stash_nb;           // its buckets (0 = empty)
    int32_t k;
    const uint32_t *postings;  // (cid, pos) pairs
    int64_t n_postings;
    const uint64_t *ref2;  // 2-bit MSB-first, 32 bases / word (+REF_PAD_WORDS zero words)
    const uint64_t *refN;  // N flags in the same geometry (low bit of the group)
    int32_t ref_has_n;
    const uint64_t *contig_off;  // n_contigs + 1
    int64_t n_contigs;
    int64_t n_bases;
};

// ------------------------------------------------------- integer hashing
// Thomas Wang 64-bit mix (the seed hash of gpu-common/src/gpu_kernels.mojo:128-137).
__host__ __device__ inline uint64_t wang_hash(uint64_t key) {
    key = (~key) + (key << 21);
    key = key ^ (key >> 24);
    key = (key + (key << 3)) + (key << 8);
    key = key ^ (key >> 14);
    key = (key + (key << 2)) + (key << 4);
    key = key ^ (key >> 28);
    key = key + (key << 31);
    return key;
}

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// SplitMix64 output mix (Steele, Lea and Flood, OOPSLA 2014): the second,
// independent bucket choice.
__host__ __device__ inline uint64_t splitmix_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t bucket_of(uint64_t key, uint64_t n_buckets) {
    return mulhi64(wang_hash(key), n_buckets);
}

// Second bucket of a key (the next bucket when both choices coincide).
__host__ __device__ inline uint64_t bucket2_of(uint64_t key, uint64_t n_buckets, uint64_t b1) {
    uint64_t b = mulhi64(splitmix_mix(key), n_buckets);
    if (b == b1) b = (b + 1 == n_buckets) ? 0 : b + 1;
    return b;
}

// ------------------------------------------------------ 2-bit word helpers
// 32 bases starting at base offset `g` of a word array (MSB-first groups).
__device__ inline uint64_t window32(const uint64_t *w, int64_t g) {
    int64_t wi = g >> 5;
    int sh = (int)(g & 31) * 2;
    uint64_t a = w[wi];
    if (sh == 0) return a;
    return (a << sh) | (w[wi + 1] >> (64 - sh));
}

// Groups of word a shifted left by sh bits (even, 0..62), filled from word b.
__device__ inline uint64_t funnel64(uint64_t a, uint64_t b, int sh) {
    const uint64_t lo = b >> ((64 - sh) & 63);
    return (a << sh) | (sh ? lo : 0ull);
}

// Reverse the order of the 32 two-bit groups of a word.
__device__ inline uint64_t reverse_groups(uint64_t x) {
    uint64_t y = __builtin_bitreverse64(x);
    return ((y >> 1) & LOW2) | ((y & LOW2) << 1);
}

// Mask with the low flag bit set for groups j < n (group j at bit 62-2j).
__device__ inline uint64_t first_groups_mask(int n) {
    if (n <= 0) return 0;
    if (n >= 32) return LOW2;
    return LOW2 & ~(~0ull >> (2 * n));
}

}  // namespace ma
