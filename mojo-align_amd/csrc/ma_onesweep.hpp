// One-sweep stable LSD radix sort for the sort + markdup path.
//
// Keys are 64-bit (one array, "lo") or 128-bit (hi, lo; order (hi, lo)); the
// value of every record is its original index (u32), so the sorted values are
// the permutation. Passes 0..7 take byte p of lo, passes 8..15 byte p - 8 of
// hi, least significant first.
//
// Everything is decided on the device, so a whole sort is a fixed sequence of
// kernel launches with no host round trip (it can be captured in a graph):
//   os_hist  — all sixteen digit histograms in one sweep (+ a check of whether
//              lo is implied by hi: lo == 0, or hi == ~0 and lo == the index);
//   os_plan  — per pass: active (no digit holds every key), exclusive digit
//              offsets, which buffer it reads and writes, first / last active
//              pass; whether lo must move with hi ("carry");
//   os_pass  — one launch per pass: each block takes tiles in order from a
//              counter, ranks its keys stably (wave ballots, per-wave digit
//              counts), publishes its per-digit counts, looks back over the
//              preceding tiles' published counts for its global offsets
//              (decoupled look-back), stages the tile in LDS in digit order
//              one array at a time (keys, the other key word, values) and
//              writes each out in runs; inactive passes return at once;
//   os_finish — the identity order when no pass is active.
// The first active pass reads the caller's key arrays and makes the values
// (i) on the fly; the last writes the permutation straight to the caller's
// output and skips the keys unless they are wanted afterwards.
#pragma once

#include "ma_sort.hpp"

namespace ma {
namespace {

constexpr int OS_BLOCK = 256;
constexpr int OS_RADIX = 256;
constexpr int OS_PASSES = 16;
constexpr int OS_ITEMS_PLAIN = 16;  // tile of 4096 keys (staged one array at a time: 32 KB of LDS)
constexpr int OS_ITEMS_CARRY = 8;   // tile of 2048 when both key words move (16 KB)
constexpr int OS_MIN_TILE = OS_BLOCK * OS_ITEMS_CARRY;
constexpr uint64_t OS_FLAG_AGG = 1ull << 62;   // tile count published
constexpr uint64_t OS_FLAG_INCL = 2ull << 62;  // inclusive prefix published
constexpr uint64_t OS_COUNT_MASK = (1ull << 62) - 1;
// Tiles go to blocks in chunks per XCD: the launch dispatches block b to XCD
// b % 8, so block group b % OS_GROUPS takes chunks of consecutive tiles from a
// counter of its own, and one XCD's L2 sees the neighbouring runs of a digit
// that its tiles write (whole lines written back instead of partial ones).
constexpr int OS_GROUPS = 8;
constexpr uint32_t OS_NO_POS = 0xFFFFFFFFu;  // no output position (sorts hold at most 2^32 - 1 keys)
constexpr uint32_t OS_SPIN_LIMIT = 1u << 22;   // look-back wait bound (~0.1 s): then error, never a hang

struct SortPlan {
    uint32_t active[OS_PASSES];
    uint32_t src[OS_PASSES];    // buffer the pass reads (when not first)
    uint32_t first[OS_PASSES];  // first active pass: reads the caller's arrays
    uint32_t last[OS_PASSES];   // last active pass: writes the caller's output
    uint32_t ord[OS_PASSES];    // ordinal among active passes (status array parity)
    uint32_t n_active;
    uint32_t fin;               // buffer holding the sorted data (0 when no pass is active)
    uint32_t carry;             // lo moves with hi (lo digits active in a 128-bit sort)
    uint32_t keep_keys;         // the last pass writes the keys too
    uint32_t has_hi;
    uint32_t error;             // a look-back wait exceeded OS_SPIN_LIMIT
    uint32_t viol;              // lo is not implied by hi (set by os_hist)
    uint32_t fviol;             // bit p: some key's byte p is 0xFF while byte p - 1 is not, or the
                                // reverse (set by os_hist; zeroed with viol)
    uint64_t digit_off[OS_PASSES][OS_RADIX];
};

struct OsBufs {
    const uint64_t *in_lo, *in_hi;  // caller's keys (in_hi null for 64-bit sorts)
    uint64_t *lo[2], *hi[2];
    uint32_t *v[2];
    uint32_t *out_v;  // permutation destination (null: v[fin])
    uint64_t *status; // 2 x tiles x 256 (cleared before the sort)
    uint32_t *tile_ctr;  // OS_GROUPS counters per pass (cleared by os_plan)
    int64_t status_stride;  // entries per status array
};

__device__ __forceinline__ uint32_t os_digit(uint64_t k, int shift) { return (uint32_t)(k >> shift) & 255u; }

// Histograms of the digits of lo (passes 0..7) and hi (8..15); a digit that is
// the same on every valid lane of a wave is counted with one LDS add. A wave
// takes OS_HIST_U chunks of 64 keys per step, their loads issued together.
constexpr int OS_HIST_U = 4;

// Bit j set where byte j of x is 0xFF (exact zero-byte flags of ~x, gathered
// into the low byte: the flag of byte j lands on bit 56 + j of the product).
__device__ __forceinline__ uint32_t ff_bytes(uint64_t x) {
    const uint64_t t = ~x;
    const uint64_t z = ~(((t & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | t | 0x7F7F7F7F7F7F7F7Full);
    return (uint32_t)(((z >> 7) * 0x0102040810204080ull) >> 56);
}

__global__ void __launch_bounds__(OS_BLOCK) os_hist_kernel(int64_t n, const uint64_t *__restrict__ lo,
                                                            const uint64_t *__restrict__ hi,
                                                            unsigned long long *__restrict__ hist,
                                                            SortPlan *__restrict__ plan) {
    __shared__ uint32_t h[OS_PASSES * OS_RADIX];
    for (int i = threadIdx.x; i < OS_PASSES * OS_RADIX; i += OS_BLOCK) h[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int np = hi ? 16 : 8;
    uint32_t viol = 0, fv = 0;
    for (int64_t base = ((int64_t)blockIdx.x * OS_BLOCK + (threadIdx.x & ~63)) * OS_HIST_U; base < n;
         base += (int64_t)gridDim.x * OS_BLOCK * OS_HIST_U) {
        uint64_t a[OS_HIST_U], b[OS_HIST_U];
#pragma unroll
        for (int u = 0; u < OS_HIST_U; u++) {
            const int64_t i = base + u * 64 + lane;
            a[u] = i < n ? lo[i] : 0ull;
            b[u] = (i < n && hi) ? hi[i] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < OS_HIST_U; u++) {
            const int64_t i = base + u * 64 + lane;
            const bool valid = i < n;
            const uint64_t vmask = __ballot(valid);
            if (vmask == 0ull) continue;  // wave-uniform
            if (valid && hi && a[u] != 0ull && !(b[u] == ~0ull && a[u] == (uint64_t)i)) viol = 1;
            if (valid) {  // bit p of m: byte p (lo bytes, then hi) is all ones; a change from p - 1
                const uint32_t m = ff_bytes(a[u]) | (hi ? ff_bytes(b[u]) << 8 : 0u);
                fv |= m ^ (m << 1);
            }
            for (int p = 0; p < np; p++) {
                const uint64_t k = p < 8 ? a[u] : b[u];
                const uint32_t d = os_digit(k, 8 * (p & 7));
                const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane(d);
                const uint64_t same = __ballot(valid && d == d0);
                if (same == vmask) {
                    if (lane == 0) atomicAdd(&h[p * OS_RADIX + d0], (uint32_t)__popcll(vmask));
                } else if (valid) {
                    atomicAdd(&h[p * OS_RADIX + d], 1u);
                }
            }
        }
    }
    if (__ballot(viol)) {
        if (lane == 0) atomicOr(&plan->viol, 1u);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fv |= (uint32_t)__shfl_xor((int)fv, o, 64);
    fv &= 0xFFFEu;  // bit 0 compares byte 0 with nothing
    if (fv && lane == 0) atomicOr(&plan->fviol, fv);
    __syncthreads();
    for (int i = threadIdx.x; i < np * OS_RADIX; i += OS_BLOCK)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// One block of 256 threads (thread = digit). check_implied: a 128-bit sort
// whose lo may be implied by hi (then the lo passes are skipped and lo does
// not move). Clears the per-pass tile counters.
__global__ void __launch_bounds__(OS_BLOCK) os_plan_kernel(SortPlan *__restrict__ plan,
                                                            const unsigned long long *__restrict__ hist,
                                                            int64_t n, int has_hi, int check_implied,
                                                            int keep_keys, uint32_t *__restrict__ tile_ctr) {
    __shared__ uint64_t sh[4];
    __shared__ uint32_t act[OS_PASSES];
    const int t = threadIdx.x;
    const bool implied = has_hi && check_implied && plan->viol == 0;
    const uint32_t fviol = plan->fviol;
    for (int p = 0; p < OS_PASSES; p++) {
        const bool exists = p < 8 || has_hi;
        const uint64_t c = exists ? hist[p * OS_RADIX + t] : 0ull;
        const int full = __syncthreads_or(exists && c == (uint64_t)n);
        // a byte that is 0xFF exactly where the byte before it is, and 0 elsewhere, is a
        // non-decreasing function of that byte: the records are already in its order, and
        // the stable pass would move nothing (the unmapped sentinels' all-ones high bytes);
        // not across lo passes skipped as implied, which leave the records unsorted by them
        const int other = __syncthreads_or(exists && c != 0ull && t != 0 && t != 0xFF);
        const bool follows = p > 0 && !(p == 8 && implied) && !other && !((fviol >> p) & 1u);
        uint64_t tot;
        const uint64_t ex = block_excl_scan_256(c, sh, tot);
        plan->digit_off[p][t] = ex;
        if (t == 0) act[p] = (exists && n > 1 && !full && !(p < 8 && implied) && !follows) ? 1u : 0u;
        __syncthreads();
    }
    if (t < OS_PASSES * OS_GROUPS) tile_ctr[t] = 0u;
    if (t == 0) {
        uint32_t carry = 0;
        if (has_hi)
            for (int p = 0; p < 8; p++) carry |= act[p];
        int last = -1, cnt = 0;
        uint32_t cur = 0;
        for (int p = 0; p < OS_PASSES; p++) {
            plan->active[p] = act[p];
            plan->first[p] = 0;
            plan->last[p] = 0;
            plan->src[p] = cur;
            plan->ord[p] = (uint32_t)cnt;
            if (!act[p]) continue;
            if (cnt == 0) {
                plan->first[p] = 1;
                cur = 0;  // the first pass writes buffer 0
            } else {
                plan->src[p] = cur;
                cur ^= 1u;
            }
            cnt++;
            last = p;
        }
        if (last >= 0) plan->last[last] = 1;
        plan->n_active = (uint32_t)cnt;
        plan->fin = cur;
        plan->carry = carry;
        plan->keep_keys = (uint32_t)keep_keys;
        plan->has_hi = (uint32_t)has_hi;
        plan->error = 0u;
    }
}

// One pass. CARRY: both key words move (the other word is staged too).
template <bool CARRY, int ITEMS>
__global__ void __launch_bounds__(OS_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) os_pass_kernel(SortPlan *__restrict__ plan, int p, int64_t n,
                                                            OsBufs B) {
    constexpr int TILE = OS_BLOCK * ITEMS;
    constexpr int WAVE_KEYS = 64 * ITEMS;
    // The tile in digit order, one array at a time (keys, the other key word,
    // values) through one buffer: 32 KB at 4096 keys, so four blocks fit a CU.
    __shared__ uint64_t s_buf[TILE];
    __shared__ uint32_t s_cnt[4][OS_RADIX];  // per-wave digit counts, then per-wave offsets
    // global position of LDS slot 0 of digit d, mod 2^32 (positions are < n <= 2^32 - 1)
    __shared__ uint32_t s_gb[OS_RADIX];
    __shared__ uint64_t s_scan[4];
    __shared__ int64_t s_tile;

    if (!plan->active[p] || (plan->carry != 0) != CARRY) return;
    const bool hi_pass = p >= 8;
    const int shift = 8 * (p & 7);
    const uint32_t src = plan->src[p], dst = plan->first[p] ? 0u : (src ^ 1u);
    const bool first = plan->first[p] != 0, last = plan->last[p] != 0;
    const bool write_keys = !last || plan->keep_keys != 0;
    const uint64_t *Din = first ? (hi_pass ? B.in_hi : B.in_lo) : (hi_pass ? B.hi[src] : B.lo[src]);
    const uint64_t *Oin = first ? (hi_pass ? B.in_lo : B.in_hi) : (hi_pass ? B.lo[src] : B.hi[src]);
    uint64_t *Dout = hi_pass ? B.hi[dst] : B.lo[dst];
    uint64_t *Oout = hi_pass ? B.lo[dst] : B.hi[dst];
    const uint32_t *Vin = B.v[src];
    uint32_t *Vout = (last && B.out_v) ? B.out_v : B.v[dst];
    uint64_t *st_cur = B.status + (int64_t)(plan->ord[p] & 1u) * B.status_stride;
    uint64_t *st_next = B.status + (int64_t)((plan->ord[p] + 1u) & 1u) * B.status_stride;
    const uint64_t *gdo = plan->digit_off[p];

    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const uint64_t lt = lanemask_lt(lane);
    const int64_t ntiles = (n + TILE - 1) / TILE;
    // block group and its chunk of consecutive tiles (every group has a block:
    // groups = min(OS_GROUPS, grid)); a group's k-th tile is increasing in k
    const int groups = (int)gridDim.x < OS_GROUPS ? (int)gridDim.x : OS_GROUPS;
    const int grp = (int)blockIdx.x % groups;
    const int64_t chunk = (int64_t)gridDim.x / groups;
    uint32_t *ctr = B.tile_ctr + p * OS_GROUPS + grp;

    while (true) {
        if (t == 0) {
            const int64_t k = (int64_t)atomicAdd(ctr, 1u);
            s_tile = (k / chunk) * chunk * groups + (int64_t)grp * chunk + k % chunk;
        }
        for (int i = t; i < 4 * OS_RADIX; i += OS_BLOCK) (&s_cnt[0][0])[i] = 0u;
        __syncthreads();
        const int64_t tile = s_tile;
        if (tile >= ntiles) break;  // block-uniform
        const int64_t tbase = tile * TILE;
        const int64_t wbase = tbase + (int64_t)w * WAVE_KEYS;
        // keys of this wave's part of the tile: item r of a lane is valid when r * 64 + lane < nv
        const int nv = (int)(n - wbase < 0 ? 0 : (n - wbase < WAVE_KEYS ? n - wbase : WAVE_KEYS));

        // rank: rounds of 64 consecutive keys per wave, lanes in order
        uint64_t key[ITEMS];
        uint64_t oth[CARRY ? ITEMS : 1];
        uint32_t val[ITEMS], rnk[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const bool valid = r * 64 + lane < nv;
            key[r] = valid ? Din[wbase + r * 64 + lane] : 0ull;
            if (CARRY) oth[CARRY ? r : 0] = valid ? Oin[wbase + r * 64 + lane] : 0ull;
        }
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const bool valid = r * 64 + lane < nv;
            const uint32_t d = os_digit(key[r], shift);
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const uint64_t m = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? m : ~m;
            }
            uint32_t base = 0;
            if (valid) base = s_cnt[w][d];
            __builtin_amdgcn_wave_barrier();
            __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (valid && (peers & lt) == 0ull) s_cnt[w][d] = base + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            rnk[r] = base + (uint32_t)__popcll(peers & lt);
        }
        __syncthreads();
        // per-digit: wave offsets, tile count, tile-local digit offsets
        const uint32_t c0 = s_cnt[0][t], c1 = s_cnt[1][t], c2 = s_cnt[2][t], c3 = s_cnt[3][t];
        const uint32_t tcnt = c0 + c1 + c2 + c3;
        uint64_t tot;
        const uint32_t dex = (uint32_t)block_excl_scan_256(tcnt, s_scan, tot);
        s_cnt[0][t] = dex;
        s_cnt[1][t] = dex + c0;
        s_cnt[2][t] = dex + c0 + c1;
        s_cnt[3][t] = dex + c0 + c1 + c2;
        // publish this tile's count of digit t
        uint64_t *my = st_cur + tile * OS_RADIX + t;
        if (tile == 0) {
            __hip_atomic_store(my, OS_FLAG_INCL | (uint64_t)tcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(my, OS_FLAG_AGG | (uint64_t)tcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // look back for the preceding tiles' counts of digit t
        uint64_t prefix = 0;
        if (tile > 0) {
            int64_t q = tile - 1;
            uint32_t spins = 0;
            while (true) {
                const uint64_t e = __hip_atomic_load(st_cur + q * OS_RADIX + t, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t f = e & ~OS_COUNT_MASK;
                if (f == 0ull) {
                    if (++spins > OS_SPIN_LIMIT) {
                        atomicOr(&plan->error, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += e & OS_COUNT_MASK;
                if (f == OS_FLAG_INCL || q == 0) break;
                q--;
            }
            __hip_atomic_store(my, OS_FLAG_INCL | (prefix + tcnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the next pass's status entries of this tile start cleared
        st_next[tile * OS_RADIX + t] = 0ull;
        s_gb[t] = (uint32_t)(gdo[t] + prefix - (uint64_t)dex);
        __syncthreads();
        // stage the keys in digit order (rnk becomes the LDS slot)
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            if (r * 64 + lane < nv) {
                rnk[r] += s_cnt[w][os_digit(key[r], shift)];
                s_buf[rnk[r]] = key[r];
            }
        }
        __syncthreads();
        // the values load while the keys are written (fewer registers held
        // through the ranking and the look-back)
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const int li = r * 64 + lane;
            val[r] = li < nv ? (first ? (uint32_t)(wbase + li) : Vin[wbase + li]) : 0u;
        }
        // write out: consecutive LDS slots of a digit go to consecutive positions;
        // slot j = t + 256 r of this thread goes to gpos[r] (OS_NO_POS: nothing to
        // write, which also holds after a look-back timeout, plan->error; n <= 2^32 - 1
        // keeps positions below it)
        const int tn = (int)((n - tbase) < TILE ? (n - tbase) : TILE);
        uint32_t gpos[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const int j = t + r * OS_BLOCK;
            gpos[r] = OS_NO_POS;
            if (j < tn) {
                const uint64_t k = s_buf[j];
                const uint32_t gp = s_gb[os_digit(k, shift)] + (uint32_t)j;
                if ((uint64_t)gp < (uint64_t)n) {
                    gpos[r] = gp;
                    if (write_keys) Dout[gp] = k;
                }
            }
            // bound the loads in flight (registers), four slots at a time
            if ((r & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        if (CARRY) {
#pragma unroll
            for (int r = 0; r < ITEMS; r++)
                if (r * 64 + lane < nv) s_buf[rnk[r]] = oth[CARRY ? r : 0];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < ITEMS; r++)
                if (write_keys && gpos[r] != OS_NO_POS) Oout[gpos[r]] = s_buf[t + r * OS_BLOCK];
            __syncthreads();
        }
        uint32_t *s_val = reinterpret_cast<uint32_t *>(s_buf);
#pragma unroll
        for (int r = 0; r < ITEMS; r++)
            if (r * 64 + lane < nv) s_val[rnk[r]] = val[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ITEMS; r++)
            if (gpos[r] != OS_NO_POS) Vout[gpos[r]] = s_val[t + r * OS_BLOCK];
        __syncthreads();
    }
}

// No active pass (n <= 1, or every key equal): the identity order; the keys,
// when wanted afterwards, copied into buffer 0.
__global__ void os_finish_kernel(const SortPlan *__restrict__ plan, int64_t n, OsBufs B) {
    if (plan->n_active != 0) return;
    uint32_t *V = B.out_v ? B.out_v : B.v[0];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        V[i] = (uint32_t)i;
        if (plan->keep_keys) {
            B.lo[0][i] = B.in_lo[i];
            if (B.in_hi) B.hi[0][i] = B.in_hi[i];
        }
    }
}

struct OsScratch {
    uint64_t *lo[2], *hi[2];
    uint32_t *v[2];
    uint64_t *status;
    int64_t status_stride;
    unsigned long long *hist;
    SortPlan *plan;
    uint32_t *tile_ctr;
};

inline int64_t os_status_entries(int64_t n) { return ((n + OS_MIN_TILE - 1) / OS_MIN_TILE + 1) * OS_RADIX; }

inline int os_grid(int64_t ntiles) {
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu <= 0)
            n_cu = 256;
    }
    const int64_t cap = (int64_t)n_cu * 4;
    return (int)(ntiles < cap ? (ntiles < 1 ? 1 : ntiles) : cap);
}

// Stable sort of n records by lo (in_hi null) or (in_hi, in_lo); out_v
// receives the permutation (null: s.v[plan->fin]); keep_keys leaves the sorted
// key words in s.lo / s.hi [plan->fin]; check_implied skips the lo passes
// when lo is implied by hi (see os_hist_kernel). Kernels only, on st.
inline int os_sort(hipStream_t st, int64_t n, const uint64_t *in_lo, const uint64_t *in_hi, uint32_t *out_v,
                   int keep_keys, int check_implied, const OsScratch &s) {
    MA_TRY(fill_u64(st, s.status, 2 * s.status_stride, 0ull));
    MA_TRY(fill_u64(st, (uint64_t *)s.hist, OS_PASSES * OS_RADIX, 0ull));
    MA_TRY(fill_u64(st, (uint64_t *)&s.plan->viol, 1, 0ull));  // viol and fviol
    if (n > 0) {
        os_hist_kernel<<<grid_for(n, OS_BLOCK, 2048), OS_BLOCK, 0, st>>>(n, in_lo, in_hi, s.hist, s.plan);
        MA_LAUNCHED();
    }
    os_plan_kernel<<<1, OS_BLOCK, 0, st>>>(s.plan, s.hist, n, in_hi ? 1 : 0, check_implied, keep_keys, s.tile_ctr);
    MA_LAUNCHED();
    OsBufs B{in_lo, in_hi, {s.lo[0], s.lo[1]}, {s.hi[0], s.hi[1]}, {s.v[0], s.v[1]}, out_v,
             s.status, s.tile_ctr, s.status_stride};
    const int np = in_hi ? OS_PASSES : 8;
    // tile of the passes that move one key word: 4096 keys (16 per thread) or 2048 (8), as tuned
    const int plain_items = tune_get(MA_TUNE_SORT_TILE_ITEMS) == 8 ? 8 : OS_ITEMS_PLAIN;
    const int g_plain = os_grid((n + OS_BLOCK * plain_items - 1) / (OS_BLOCK * plain_items));
    const int g_carry = os_grid((n + OS_BLOCK * OS_ITEMS_CARRY - 1) / (OS_BLOCK * OS_ITEMS_CARRY));
    for (int p = 0; p < np && n > 1; p++) {
        if (in_hi) {
            os_pass_kernel<true, OS_ITEMS_CARRY><<<g_carry, OS_BLOCK, 0, st>>>(s.plan, p, n, B);
            MA_LAUNCHED();
        }
        if (!in_hi || p >= 8) {
            if (plain_items == 8)
                os_pass_kernel<false, 8><<<g_plain, OS_BLOCK, 0, st>>>(s.plan, p, n, B);
            else
                os_pass_kernel<false, OS_ITEMS_PLAIN><<<g_plain, OS_BLOCK, 0, st>>>(s.plan, p, n, B);
            MA_LAUNCHED();
        }
    }
    if (n > 0) {
        os_finish_kernel<<<grid_for(n), 256, 0, st>>>(s.plan, n, B);
        MA_LAUNCHED();
    }
    return MA_OK;
}

}  // namespace
}  // namespace ma
