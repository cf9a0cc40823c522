// Device-wide primitives: stable LSD radix sort (8-bit digits, u64 keys with
// u32/u64 values), exclusive scan of u32 flags, fills. Wave64 ballots rank the
// keys of a tile; digit-major block offsets keep the sort stable across tiles.
#pragma once

#include "ma_common.hpp"

namespace ma {
// Internal linkage: every translation unit gets its own copy of these kernels.
namespace {

constexpr int RS_BLOCK = 256;  // 4 waves
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;

__device__ inline uint64_t lanemask_lt(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// Exclusive scan of one value per thread over a 256-thread block; returns the
// block total through `total`. Uses `sh` (>= 4 entries of LDS).
__device__ inline uint64_t block_excl_scan_256(uint64_t v, uint64_t *sh, uint64_t &total) {
    int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    uint64_t s0 = sh[0], s1 = sh[1], s2 = sh[2], s3 = sh[3];
    uint64_t pre = (w > 0 ? s0 : 0) + (w > 1 ? s1 : 0) + (w > 2 ? s2 : 0);
    total = s0 + s1 + s2 + s3;
    __syncthreads();
    return pre + x - v;
}

__global__ void fill_u64_kernel(uint64_t *p, int64_t n, uint64_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void iota_u32_kernel(uint32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

inline int grid_for(int64_t n, int block = 256, int cap = 65536) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

inline int fill_u64(hipStream_t st, uint64_t *p, int64_t n, uint64_t v) {
    if (n <= 0) return MA_OK;
    fill_u64_kernel<<<grid_for(n), 256, 0, st>>>(p, n, v);
    MA_LAUNCHED();
    return MA_OK;
}

// ------------------------------------------------------------ radix sort

// All eight byte-digit histograms of the keys in one sweep.
__global__ void rs_hist8_kernel(const uint64_t *__restrict__ keys, int64_t n,
                                unsigned long long *__restrict__ hist8) {
    __shared__ unsigned int h[8 * 256];
    for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t k = keys[i];
#pragma unroll
        for (int d = 0; d < 8; d++) atomicAdd(&h[d * 256 + ((k >> (8 * d)) & 255)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) {
        unsigned v = h[i];
        if (v) atomicAdd(&hist8[i], (unsigned long long)v);
    }
}

// Per-tile digit counts, digit-major: blockhist[d * nblocks + b].
__global__ void __launch_bounds__(RS_BLOCK) rs_upsweep_kernel(const uint64_t *__restrict__ keys,
                                                               int64_t n, int shift,
                                                               uint32_t *__restrict__ blockhist,
                                                               int64_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int i = 0; i < RS_ITEMS; i++) {
        int64_t idx = base + i * RS_BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255], 1u);
    }
    __syncthreads();
    blockhist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Block d turns the counts of digit d into absolute output offsets.
__global__ void __launch_bounds__(RS_BLOCK) rs_digit_scan_kernel(
    const uint32_t *__restrict__ blockhist, uint64_t *__restrict__ blockoff, int64_t nblocks,
    const unsigned long long *__restrict__ hist_pass) {
    __shared__ uint64_t sh[4];
    __shared__ uint64_t base_s;
    int d = blockIdx.x;
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int j = 0; j < d; j++) s += hist_pass[j];
        base_s = s;
    }
    __syncthreads();
    uint64_t run = base_s;
    for (int64_t c = 0; c < nblocks; c += RS_BLOCK) {
        int64_t i = c + threadIdx.x;
        uint64_t v = i < nblocks ? blockhist[(int64_t)d * nblocks + i] : 0;
        uint64_t tot;
        uint64_t ex = block_excl_scan_256(v, sh, tot);
        if (i < nblocks) blockoff[(int64_t)d * nblocks + i] = run + ex;
        run += tot;
    }
}

// Stable scatter of one tile: rounds of 256 keys in index order; within a
// round, waves in order; within a wave, lanes ranked by ballot peer masks.
template <class V>
__global__ void __launch_bounds__(RS_BLOCK) rs_scatter_kernel(
    const uint64_t *__restrict__ kin, const V *__restrict__ vin, uint64_t *__restrict__ kout,
    V *__restrict__ vout, int64_t n, int shift, const uint64_t *__restrict__ blockoff,
    int64_t nblocks) {
    __shared__ uint32_t cnt[4][256];
    __shared__ uint32_t run[256];
    __shared__ uint64_t boff[256];
    int t = threadIdx.x, w = t >> 6, lane = t & 63;
    boff[t] = blockoff[(int64_t)t * nblocks + blockIdx.x];
    run[t] = 0;
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    uint64_t lt = lanemask_lt(lane);
    for (int i = 0; i < RS_ITEMS; i++) {
        int64_t idx = base + (int64_t)i * RS_BLOCK + t;
        if (base + (int64_t)i * RS_BLOCK >= n) break;  // block-uniform
        cnt[0][t] = 0;
        cnt[1][t] = 0;
        cnt[2][t] = 0;
        cnt[3][t] = 0;
        __syncthreads();
        bool valid = idx < n;
        uint64_t k = valid ? kin[idx] : 0ull;
        V v = valid ? vin[idx] : V(0);
        uint32_t d = (uint32_t)(k >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) cnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t r0 = run[t];
        uint32_t c0 = cnt[0][t], c1 = cnt[1][t], c2 = cnt[2][t], c3 = cnt[3][t];
        __syncthreads();
        cnt[0][t] = r0;
        cnt[1][t] = r0 + c0;
        cnt[2][t] = r0 + c0 + c1;
        cnt[3][t] = r0 + c0 + c1 + c2;
        run[t] = r0 + c0 + c1 + c2 + c3;
        __syncthreads();
        if (valid) {
            uint64_t pos = boff[d] + cnt[w][d] + rank;
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
    }
}

struct RadixScratch {
    uint32_t *blockhist;        // 256 * nblocks
    uint64_t *blockoff;         // 256 * nblocks
    unsigned long long *hist8;  // 8 * 256 (device)
};

inline int64_t rs_nblocks(int64_t n) { return (n + RS_TILE - 1) / RS_TILE; }

inline size_t rs_scratch_bytes(int64_t n) {
    int64_t nb = rs_nblocks(n);
    return (size_t)(256 * nb * 4 + 256 * nb * 8 + 8 * 256 * 8 + 1024);
}

// Sort (keys, vals) stably by the low `key_bits` bits. Passes whose digit is
// the same for every key are skipped (known from one histogram sweep; one
// small device->host read). On return `keys`/`vals` point at whichever
// buffer of each pair holds the result.
template <class V>
int radix_sort_pairs(hipStream_t st, int64_t n, uint64_t *&keys, V *&vals, uint64_t *&kalt,
                     V *&valt, const RadixScratch &s, int key_bits = 64) {
    if (n <= 1) return MA_OK;
    unsigned long long h_hist8[8 * 256];
    MA_TRY(fill_u64(st, (uint64_t *)s.hist8, 8 * 256, 0));
    rs_hist8_kernel<<<grid_for(n, 256, 2048), 256, 0, st>>>(keys, n, s.hist8);
    MA_LAUNCHED();
    MA_HIP(hipMemcpyAsync(h_hist8, s.hist8, sizeof(h_hist8), hipMemcpyDeviceToHost, st));
    MA_HIP(hipStreamSynchronize(st));
    int64_t nb = rs_nblocks(n);
    for (int p = 0; p < 8 && p * 8 < key_bits; p++) {
        bool trivial = false;
        for (int d = 0; d < 256; d++)
            if (h_hist8[p * 256 + d] == (unsigned long long)n) trivial = true;
        if (trivial) continue;
        rs_upsweep_kernel<<<(unsigned)nb, RS_BLOCK, 0, st>>>(keys, n, p * 8, s.blockhist, nb);
        MA_LAUNCHED();
        rs_digit_scan_kernel<<<256, RS_BLOCK, 0, st>>>(s.blockhist, s.blockoff, nb, s.hist8 + p * 256);
        MA_LAUNCHED();
        rs_scatter_kernel<V><<<(unsigned)nb, RS_BLOCK, 0, st>>>(keys, vals, kalt, valt, n, p * 8,
                                                                 s.blockoff, nb);
        MA_LAUNCHED();
        uint64_t *tk = keys;
        keys = kalt;
        kalt = tk;
        V *tv = vals;
        vals = valt;
        valt = tv;
    }
    return MA_OK;
}

// -------------------------------------------------------------- scan

__global__ void __launch_bounds__(RS_BLOCK) scan_partials_kernel(const uint32_t *__restrict__ in,
                                                                 int64_t n,
                                                                 uint64_t *__restrict__ partial) {
    __shared__ uint64_t sh[4];
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    uint64_t s = 0;
    for (int i = 0; i < RS_ITEMS; i++) {
        int64_t idx = base + (int64_t)i * RS_BLOCK + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    uint64_t tot;
    (void)block_excl_scan_256(s, sh, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(RS_BLOCK) scan_top_kernel(uint64_t *__restrict__ partial,
                                                            int64_t nb, uint64_t *__restrict__ total) {
    __shared__ uint64_t sh[4];
    uint64_t run = 0;
    for (int64_t c = 0; c < nb; c += RS_BLOCK) {
        int64_t i = c + threadIdx.x;
        uint64_t v = i < nb ? partial[i] : 0;
        uint64_t tot;
        uint64_t ex = block_excl_scan_256(v, sh, tot);
        if (i < nb) partial[i] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0 && total) *total = run;
}

__global__ void __launch_bounds__(RS_BLOCK) scan_final_kernel(const uint32_t *__restrict__ in,
                                                              int64_t n,
                                                              const uint64_t *__restrict__ partial,
                                                              uint64_t *__restrict__ out) {
    __shared__ uint64_t sh[4];
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    uint64_t run = partial[blockIdx.x];
    for (int i = 0; i < RS_ITEMS; i++) {
        int64_t r0 = base + (int64_t)i * RS_BLOCK;
        if (r0 >= n) break;  // block-uniform
        int64_t idx = r0 + threadIdx.x;
        uint64_t v = idx < n ? in[idx] : 0;
        uint64_t tot;
        uint64_t ex = block_excl_scan_256(v, sh, tot);
        if (idx < n) out[idx] = run + ex;
        run += tot;
    }
}

inline size_t scan_scratch_bytes(int64_t n) { return (size_t)(rs_nblocks(n) * 8 + 256); }

// out[i] = sum of in[0..i); *d_total (device, may be null) = sum of all.
inline int exclusive_scan_u32(hipStream_t st, int64_t n, const uint32_t *in, uint64_t *out,
                              uint64_t *partial, uint64_t *d_total) {
    if (n <= 0) return MA_OK;
    int64_t nb = rs_nblocks(n);
    scan_partials_kernel<<<(unsigned)nb, RS_BLOCK, 0, st>>>(in, n, partial);
    MA_LAUNCHED();
    scan_top_kernel<<<1, RS_BLOCK, 0, st>>>(partial, nb, d_total);
    MA_LAUNCHED();
    scan_final_kernel<<<(unsigned)nb, RS_BLOCK, 0, st>>>(in, n, partial, out);
    MA_LAUNCHED();
    return MA_OK;
}

}  // namespace
}  // namespace ma
