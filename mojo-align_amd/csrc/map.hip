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
//     vote table, KEEP picks) lives in LDS (~8.5 KB);
//   * reads and reference are 2-bit packed (32 bases per u64, MSB-first) with
//     parallel N-flag words, so one lane scores a whole read against a
//     reference window with ~2W XOR/popcount ops (W = read words);
//   * lanes run in parallel over seeds (encode + one 64-byte hash-bucket probe
//     each), over the postings of one seed (vote / rescue / the exact tally's
//     first high-occ seed, which decides the tally in closed form) and
//     over (pick, shift) extend candidates; ballots and prefix counts keep the
//     reference's first-seen table order, and first-minimum reductions keep its
//     sequential best-hit choice (picks are in non-increasing vote order, so
//     the vote tie-break of the extend never fires);
//   * the rare vote-table overflow (more than 256 distinct loci) runs the
//     reference's first-min replacement one posting at a time.
// No host round trip between passes: FW pass 0 / pass 1, the need_rc test, RC
// passes and the FW/RC merge all run inside the kernel.
#include <string.h>

#include <mutex>

#include "ma_index.hpp"
#include "ma_sort.hpp"

using namespace ma;

namespace {

constexpr int TOP = 256;  // vote table (TOP, linear_gpu_locate.mojo:812)
constexpr int KEEP = 16;  // extend picks (KEEP, :813)
constexpr int VHASH_BITS = 9;  // vote hash: 512 slots for at most TOP keys
constexpr int VHASH = 1 << VHASH_BITS;
constexpr int HX = 48;    // exact tally entries (HX, :1249); the tally's closed form needs HX >= 2
static_assert(HX >= 2, "the closed-form tally tells one entry from several");
constexpr int MAXS = 64;  // seeds per read (one per lane)
constexpr int WPB = 4;    // waves per block
constexpr int BLOCK = WPB * 64;
constexpr int MAX_HIGH_TRIED = 96;  // :1277-1281

constexpr int PBUF = 256;  // postings prefetched into LDS per wave
constexpr int RESIDENT_WAVES_PER_CU = 16;  // 4 waves/SIMD x 4 SIMDs at the kernels' register budget

struct Knobs {
    int32_t k, stride, max_occ, vote_occ, vote_fast, max_diff, max_soft, do_rc, max_indel;
    int32_t rc_out, fw_out;  // the caller wants the per-strand RC / FW records too
};

// ind packs the single indel of a hit: iop (bits 0-1: 0 none, 1 insertion,
// 2 deletion) | ilen << 2 | iat << 8 (query bases before the indel).
struct Rec {
    int32_t cid;
    uint32_t pos, mapq;
    int32_t flag;
    uint32_t sl, sr, nm, ind;
};

__host__ __device__ constexpr uint32_t pack_ind(uint32_t iop, uint32_t ilen, uint32_t iat) {
    return iop | ilen << 2 | iat << 8;
}

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
    int64_t n;  // reads (with n_dev: the capacity of the arrays)
    int32_t max_len;
    int32_t encoding;
    int32_t bs_r1, bs_r2;
    int64_t n_r1;
    // Read from device memory by the kernel when set (a launch whose shape
    // does not depend on a host-known count): n_dev the number of reads,
    // base_dev the first read's row in bases / lens / out_idx; out_idx[r] the
    // slot read r's records go to (null: r).
    const uint32_t *n_dev;
    const uint32_t *base_dev;
    const uint32_t *out_idx;
};

// Arguments of the forward host loop's kernel (fwd_map_kernel).
struct FwdArgs {
    const uint8_t *bases;  // ASCII rows
    int64_t stride;
    const uint32_t *lens;
    int64_t n, n_r1;
    int32_t max_len, bs_r1, bs_r2;
    int32_t k, seed_stride, max_occ, max_diff, max_soft;
    int32_t count_only;               // 1: only the largest count of voting postings of a read
    unsigned long long *max_total;    // ... into here
    uint32_t *tags;                   // per-wave tables of 1 << cap_log2 entries
    uint32_t *cnts;
    uint64_t *keys;
    int32_t cap_log2;
    uint32_t tag_base;                // read r's entries carry tag_base + r + 1
    unsigned int *counter;            // work claims
    int32_t claim;
};

#define DEV __device__ __forceinline__

DEV Rec unmapped_rec() { return Rec{-1, 0u, 0u, 4, 0u, 0u, 0u, 0u}; }

DEV Rec sel_rec(bool c, const Rec &a, const Rec &b) {
    Rec r;
    r.cid = c ? a.cid : b.cid;
    r.pos = c ? a.pos : b.pos;
    r.mapq = c ? a.mapq : b.mapq;
    r.flag = c ? a.flag : b.flag;
    r.sl = c ? a.sl : b.sl;
    r.sr = c ? a.sr : b.sr;
    r.nm = c ? a.nm : b.nm;
    r.ind = c ? a.ind : b.ind;
    return r;
}

DEV int lane_id() { return (int)(threadIdx.x & 63); }

// Wave-uniform values read from LDS live in scalar registers. The lane-read
// builtins return int: each 32-bit half goes through uint32_t before it is
// widened, or a low word with bit 31 set would sign-extend over the high word
// (a posting start >= 2^31, the inline-posting marker).
DEV uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane(v); }
DEV int unii(int v) { return (int)__builtin_amdgcn_readfirstlane((uint32_t)v); }
DEV uint64_t uni64(uint64_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}

// Orders LDS traffic between lanes of one wave (lockstep lanes, in-order LDS
// queue per wave): compiler barrier + wait for outstanding LDS operations.
DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

DEV uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint64_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

DEV uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

DEV uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

DEV uint64_t readlane64(uint64_t v, int lane) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((uint32_t)v, lane);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}

// Multiset top-2 (b1 >= b2), one value at a time, with selects only.
DEV void top2_push(int v, int &b1, int &b2) {
    int nb2 = v > b1 ? b1 : (v > b2 ? v : b2);
    int nb1 = v > b1 ? v : b1;
    b1 = nb1;
    b2 = nb2;
}

DEV void wave_top2(int &b1, int &b2) {
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

// 32 bases at base offset p of a zero-padded LDS word array.
DEV uint64_t lds_window(const uint64_t *w, int p) {
    int wi = p >> 5, sh = (p & 31) * 2;
    uint64_t a = w[wi];
    if (sh == 0) return a;
    return (a << sh) | (w[wi + 1] >> (64 - sh));
}

// Both bits of the groups j < n.
DEV uint64_t groups_mask2(int n) {
    uint64_t f = first_groups_mask(n);
    return f | (f << 1);
}

struct Caps {
    int budget, clip_cap, shift_cap, indel_cap;
};

DEV Caps make_caps(const Knobs &kn, int L) {
    Caps c;
    int budget = kn.max_diff;
    if (budget < 2) budget = 2;
    if (budget > L / 2) budget = L / 2;
    int clip = kn.max_soft;
    if (clip > L / 4) clip = L / 4;
    // indel_cap = clamp(max_indel, 0, 4), shift_cap = clamp(indel_cap, 2, 6)
    // (linear_gpu_locate.mojo:1015-1024)
    int indel = kn.max_indel < 0 ? 0 : (kn.max_indel > 4 ? 4 : kn.max_indel);
    c.budget = budget;
    c.clip_cap = clip;
    c.indel_cap = indel;
    c.shift_cap = indel < 2 ? 2 : indel;
    return c;
}

DEV uint64_t load_posting(const IndexView &ix, int64_t pi) {
    // (cid, pos) u32 pair: cid in the low word
    return *(const uint64_t *)(ix.postings + 2 * pi);
}

// MAPQ 20, 40 at votes >= 3, 60 at votes >= 5; 20 at most when cost > 2 or
// the hit has an indel (linear_gpu_locate.mojo:1587-1597).
DEV Rec finish_rec(int64_t cid, int64_t pos, int votes, int cost, int sl, int sr,
                   int32_t strand_flag, uint32_t ind = 0) {
    Rec r;
    uint32_t mq = 20;
    if (votes >= 3) mq = 40;
    if (votes >= 5) mq = 60;
    if (cost > 2 && mq > 20) mq = 20;
    if ((ind & 3u) != 0 && mq > 20) mq = 20;
    r.cid = (int32_t)cid;
    r.pos = (uint32_t)pos;
    r.mapq = mq;
    r.flag = strand_flag;
    r.sl = (uint32_t)sl;
    r.sr = (uint32_t)sr;
    r.nm = (uint32_t)cost;
    r.ind = ind;
    return r;
}

DEV uint8_t ascii_code(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 255;
    }
}

// Record outputs as one table of field arrays: entry 10 * o + f is field f
// (cid pos mapq flag sl sr nm iop iat ilen) of output o (merged, FW, RC), null
// when that output or field is absent. The map kernel keeps the table in LDS
// and stores a read's records with one lane per field, so the 30 pointers hold
// no scalar registers across the kernel.
constexpr int REC_FIELDS = 10;
constexpr int REC_WORDS = 8;  // Rec's words: seven fields, then iop / ilen / iat packed
constexpr int N_REC_PTRS = 3 * REC_FIELDS;
static_assert(sizeof(Rec) == 4 * REC_WORDS, "Rec is eight 32-bit words");

DEV uint32_t *rec_field(const RecOut &o, int f) {
    if (!o.cid) return nullptr;
    uint32_t *p = (uint32_t *)o.cid;
    p = f == 1 ? o.pos : p;
    p = f == 2 ? o.mapq : p;
    p = f == 3 ? (uint32_t *)o.flag : p;
    p = f == 4 ? o.sl : p;
    p = f == 5 ? o.sr : p;
    p = f == 6 ? o.nm : p;
    p = f == 7 ? o.iop : p;
    p = f == 8 ? o.iat : p;
    p = f == 9 ? o.ilen : p;
    return p;
}

}  // namespace

namespace w4 {
namespace lds {
namespace {
constexpr int MAXW = 4;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 4;
constexpr bool COFF_IN_LDS = true;
constexpr bool INDEL = false;
#include "map_body.inc"
}  // namespace
}  // namespace lds
}  // namespace w4

namespace w4 {
namespace glb {
namespace {
constexpr int MAXW = 4;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 4;
constexpr bool COFF_IN_LDS = false;
constexpr bool INDEL = false;
#include "map_body.inc"
}  // namespace
}  // namespace glb
}  // namespace w4

namespace w5 {
namespace lds {
namespace {
constexpr int MAXW = 5;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 4;
constexpr bool COFF_IN_LDS = true;
constexpr bool INDEL = false;
#include "map_body.inc"
}  // namespace
}  // namespace lds
}  // namespace w5

namespace w5 {
namespace glb {
namespace {
constexpr int MAXW = 5;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 4;
constexpr bool COFF_IN_LDS = false;
constexpr bool INDEL = false;
#include "map_body.inc"
}  // namespace
}  // namespace glb
}  // namespace w5

namespace w8 {
namespace lds {
namespace {
constexpr int MAXW = 8;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 3;
constexpr bool COFF_IN_LDS = true;
constexpr bool INDEL = false;
#define MA_WITH_FWD  // the forward host loop's kernel too (reads up to 256 bases)
#include "map_body.inc"
#undef MA_WITH_FWD
}  // namespace
}  // namespace lds
}  // namespace w8

namespace w8 {
namespace glb {
namespace {
constexpr int MAXW = 8;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 3;
constexpr bool COFF_IN_LDS = false;
constexpr bool INDEL = false;
#define MA_WITH_FWD  // the forward host loop's kernel too (reads up to 256 bases)
#include "map_body.inc"
#undef MA_WITH_FWD
}  // namespace
}  // namespace glb
}  // namespace w8

// The single-indel branch (max_indel 1..4): its own instantiations (contig
// offsets in LDS or HBM, as for the gapless kernels), so the gapless kernels
// carry none of its code or registers; two waves per SIMD leave its extra
// window masks in registers.
namespace w4 {
namespace ind_l {
namespace {
constexpr int MAXW = 4;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 2;
constexpr bool COFF_IN_LDS = true;
constexpr bool INDEL = true;
#include "map_body.inc"
}  // namespace
}  // namespace ind_l
}  // namespace w4

namespace w4 {
namespace ind_g {
namespace {
constexpr int MAXW = 4;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 2;
constexpr bool COFF_IN_LDS = false;
constexpr bool INDEL = true;
#include "map_body.inc"
}  // namespace
}  // namespace ind_g
}  // namespace w4

namespace w5 {
namespace ind_l {
namespace {
constexpr int MAXW = 5;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 2;
constexpr bool COFF_IN_LDS = true;
constexpr bool INDEL = true;
#include "map_body.inc"
}  // namespace
}  // namespace ind_l
}  // namespace w5

namespace w5 {
namespace ind_g {
namespace {
constexpr int MAXW = 5;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 2;
constexpr bool COFF_IN_LDS = false;
constexpr bool INDEL = true;
#include "map_body.inc"
}  // namespace
}  // namespace ind_g
}  // namespace w5

namespace w8 {
namespace ind_l {
namespace {
constexpr int MAXW = 8;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 2;
constexpr bool COFF_IN_LDS = true;
constexpr bool INDEL = true;
#include "map_body.inc"
}  // namespace
}  // namespace ind_l
}  // namespace w8

namespace w8 {
namespace ind_g {
namespace {
constexpr int MAXW = 8;
constexpr int MAX_READ = MAXW * 32;
constexpr int WAVES_PER_EU = 2;
constexpr bool COFF_IN_LDS = false;
constexpr bool INDEL = true;
#include "map_body.inc"
}  // namespace
}  // namespace ind_g
}  // namespace w8

namespace {

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

}  // namespace

namespace ma {

int check_map_params(const ma_index *ix, const ma_params *p, int32_t max_len) {
    MA_REQUIRE(p, "map: null params");
    MA_REQUIRE(p->k == ix->k, "map: params.k differs from the index k");
    MA_REQUIRE(p->stride >= 1, "map: stride must be >= 1");
    MA_REQUIRE(p->max_occ >= 1 && p->max_occ < (1 << 17), "map: max_occ must be in 1..131071");
    MA_REQUIRE(max_len >= 0 && max_len <= 256, "map: reads longer than 256 bases are not supported");
    int32_t ml = max_len < p->k ? p->k : max_len;
    MA_REQUIRE((ml - p->k) / p->stride + 1 <= MAXS, "map: more than 64 seed slots per read");
    return MA_OK;
}

}  // namespace ma

// The scratch of ma_map_batch_dev_noalloc: the persistent grid's work counter
// (4 of these 256 bytes). Each device holds its own copy for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress
// returns for that device, taken once per device with the device current.
// Calls of the no-alloc entry point that overlap on one device share the
// counter and race.
constexpr int NOALLOC_SCRATCH_BYTES = 256;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned int scratch[NOALLOC_SCRATCH_BYTES / 4];

namespace {

std::mutex g_scratch_mu;
unsigned int *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch(int device, unsigned int **out) {
    MA_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, "ma_map_batch_dev_noalloc: device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        MA_HIP(hipGetDevice(&cur));
        if (cur != device) MA_HIP(hipSetDevice(device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr) {
            set_error(std::string("ma_map_batch_dev_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return MA_ERR_HIP;
        }
        g_scratch_ptr[device] = (unsigned int *)ptr;
    }
    *out = g_scratch_ptr[device];
    return MA_OK;
}

}  // namespace

namespace {

// Read-length classes of the map kernels: 0 = at most 4 read words (128
// bases), 1 = 5 words (160), 2 = 8 words (256).
constexpr int SPLIT_MIN_READS = 65536;  // smaller uneven batches map in one launch

__host__ __device__ inline int len_class(int64_t L) { return L <= 4 * 32 ? 0 : (L <= 5 * 32 ? 1 : 2); }

__global__ void class_zero_kernel(uint32_t *c) {
    if (threadIdx.x < 12) c[threadIdx.x] = 0u;
}

// Reads per length class; one atomic per class and wave.
__global__ void class_count_kernel(const uint32_t *__restrict__ lens, int64_t n, uint32_t *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); base < n;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + lane;
        const int c = i < n ? len_class(lens[i]) : -1;
        for (int k = 0; k < 3; k++) {
            const uint64_t m = __ballot(c == k);
            if (lane == 0 && m) atomicAdd(&cnt[k], (uint32_t)__popcll(m));
        }
    }
}

__global__ void class_offsets_kernel(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ off,
                                     uint32_t *__restrict__ cur) {
    off[0] = 0u;
    off[1] = cnt[0];
    off[2] = cnt[0] + cnt[1];
    cur[0] = cur[1] = cur[2] = 0u;
}

// Each read's row, length and own index into its class's region of the
// sub-batch (order within a class is free: a read's records do not depend on
// its neighbours).
__global__ void class_place_kernel(const uint8_t *__restrict__ bases, int64_t stride, const uint32_t *__restrict__ lens,
                                   int64_t n, const uint32_t *__restrict__ off, uint32_t *__restrict__ cur,
                                   uint8_t *__restrict__ sb, uint32_t *__restrict__ sl, uint32_t *__restrict__ sidx) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); base < n;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + lane;
        const uint32_t L = i < n ? lens[i] : 0u;
        const int c = i < n ? len_class(L) : -1;
        int64_t slot = -1;
        for (int k = 0; k < 3; k++) {
            const uint64_t m = __ballot(c == k);
            if (!m) continue;
            uint32_t b = 0;
            if (lane == (int)__builtin_ctzll(m)) b = atomicAdd(&cur[k], (uint32_t)__popcll(m));
            b = (uint32_t)__shfl(b, (int)__builtin_ctzll(m), 64);
            if (c == k) slot = (int64_t)off[k] + b + __popcll(m & lt);
        }
        if (slot >= 0) {
            const uint8_t *src = bases + i * stride;
            uint8_t *dst = sb + slot * stride;
            for (int64_t j = 0; j < stride; j++) dst[j] = src[j];
            sl[slot] = L;
            sidx[slot] = (uint32_t)i;
        }
    }
}

// The map kernel of a read-length class (and indel mode, contig-offset place).
int dispatch_map(int cls, bool indel, bool lds, unsigned g, hipStream_t st, const IndexView &v, const Knobs &kn,
                 const ReadsIn &in, const RecOut &mo, const RecOut &fo, const RecOut &ro, unsigned int *ctr,
                 int claim) {
    if (indel) {
        if (cls == 0) {
            if (lds) w4::ind_l::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
            else w4::ind_g::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
        } else if (cls == 1) {
            if (lds) w5::ind_l::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
            else w5::ind_g::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
        } else {
            if (lds) w8::ind_l::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
            else w8::ind_g::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
        }
    } else if (cls == 0) {
        if (lds) w4::lds::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
        else w4::glb::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
    } else if (cls == 1) {
        if (lds) w5::lds::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
        else w5::glb::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
    } else {
        if (lds) w8::lds::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
        else w8::glb::map_reads_kernel<<<g, BLOCK, 0, st>>>(v, kn, in, mo, fo, ro, ctr, claim);
    }
    MA_LAUNCHED();
    return MA_OK;
}

// Validates one batch and enqueues its map launch on st. persistent: a grid of
// at most 8 blocks per CU taking read chunks from the counter at ctr_buf (zeroed
// here by a one-thread kernel, not a memset, so the launch can be captured);
// otherwise a grid sized to the batch, one chunk per wave (ctr unused).
int map_launch(ma_index *ix, hipStream_t st, const uint8_t *d_bases, int64_t stride, const uint32_t *d_lens,
               int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2, int64_t n_r1,
               const ma_params *p, ma_recs *d_merged, ma_recs *d_fw, ma_recs *d_rc, int32_t persistent,
               int32_t chunks_per_wave, unsigned int *ctr_buf, ma::DevBuf *split_buf, const char *who) {
    MA_REQUIRE(n >= 0 && n < (1ll << 32), std::string(who) + ": n out of range");
    MA_REQUIRE(encoding == MA_ENC_ASCII || encoding == MA_ENC_2BIT, std::string(who) + ": bad encoding");
    MA_REQUIRE(encoding == MA_ENC_ASCII ? stride >= max_len : stride * 4 >= max_len,
               std::string(who) + ": stride smaller than a read row");
    MA_TRY(check_map_params(ix, p, max_len));
    if (n == 0) return MA_OK;
    RecOut mo = rec_out(d_merged);
    MA_REQUIRE(mo.cid, std::string(who) + ": merged record arrays must be non-null");
    Knobs kn{p->k, p->stride, p->max_occ, p->vote_occ, p->vote_fast, p->max_diff, p->max_soft, p->do_rc,
             p->max_indel, 0, 0};
    auto wanted = [](const ma_recs *o) {
        return o && (o->cid || o->pos || o->mapq || o->flag || o->sl || o->sr || o->nm || o->iop || o->iat ||
                     o->ilen);
    };
    kn.rc_out = wanted(d_rc) ? 1 : 0;
    kn.fw_out = wanted(d_fw) ? 1 : 0;
    ReadsIn in{d_bases, stride, d_lens, n, max_len, encoding, bs_r1, bs_r2, n_r1, nullptr, nullptr, nullptr};
    // Size the launch to the batch: claim = reads per wave chunk (at most the
    // tuned cap) so that the resident waves all get `chunks_per_wave` chunks
    // (more, smaller chunks shorten the tail of a small batch). Persistent
    // mode: at most the tuned blocks per CU loop over chunks taken from a
    // counter. Batch-sized mode: one chunk per wave, blocks retire as they
    // finish (concurrent launches interleave).
    const int64_t cpw = chunks_per_wave > 0 ? chunks_per_wave : tune_get(MA_TUNE_MAP_CHUNKS_PER_WAVE);
    const int64_t waves = (int64_t)ix->n_cu * RESIDENT_WAVES_PER_CU * cpw;
    const int64_t max_claim = tune_get(MA_TUNE_MAP_MAX_CLAIM);
    int64_t claim = n / waves;
    if (claim < 1) claim = 1;
    if (claim > max_claim) claim = max_claim;
    int64_t blocks = (n + claim * WPB - 1) / (claim * WPB);
    unsigned int *ctr = nullptr;
    if (persistent) {
        int64_t cap = (int64_t)ix->n_cu * tune_get(MA_TUNE_MAP_BLOCKS_PER_CU);
        if (blocks > cap) blocks = cap;
        ctr = ctr_buf;
    }
    MA_REQUIRE(blocks < (1ll << 31), std::string(who) + ": batch too large for one launch");
    const IndexView v = index_view(ix);
    const RecOut fo = rec_out(d_fw), ro = rec_out(d_rc);
    const bool lds = ix->n_contigs < COFF_LDS;
    const int top = len_class(max_len);
    // Uneven reads that span length classes: each class maps with its own
    // kernel (register arrays sized to its reads) from a sub-batch gathered on
    // the device; records go straight to the reads' own slots.
    const bool split = split_buf && tune_get(MA_TUNE_MAP_CLASS_SPLIT) != 0 && top > 0 && n >= SPLIT_MIN_READS &&
                       (n_r1 >= n || bs_r1 == bs_r2);
    if (!split) {
        if (ctr) {
            zero_u32_kernel<<<1, 1, 0, st>>>(ctr);
            MA_LAUNCHED();
        }
        // the read-word class of the batch's longest read picks the kernel, the
        // contig count where the contig offsets live
        return dispatch_map(top, p->max_indel > 0, lds, (unsigned)blocks, st, v, kn, in, mo, fo, ro, ctr, (int)claim);
    }
    const size_t row = (size_t)stride;
    MA_TRY(split_buf->reserve(256 + (size_t)n * (row + 8)));
    char *sb = (char *)split_buf->p;
    uint32_t *cnt = (uint32_t *)sb, *off = cnt + 4, *cur = cnt + 8;
    uint8_t *s_bases = (uint8_t *)(sb + 256);
    uint32_t *s_lens = (uint32_t *)(s_bases + (size_t)n * row);
    uint32_t *s_idx = s_lens + n;
    const int gs = grid_for(n);
    class_zero_kernel<<<1, 64, 0, st>>>(cnt);
    MA_LAUNCHED();
    class_count_kernel<<<gs, 256, 0, st>>>(d_lens, n, cnt);
    MA_LAUNCHED();
    class_offsets_kernel<<<1, 1, 0, st>>>(cnt, off, cur);
    MA_LAUNCHED();
    class_place_kernel<<<gs, 256, 0, st>>>(d_bases, stride, d_lens, n, off, cur, s_bases, s_lens, s_idx);
    MA_LAUNCHED();
    const int class_max[3] = {4 * 32, 5 * 32, 8 * 32};
    for (int c = 0; c <= top; c++) {
        ReadsIn ic{s_bases, stride, s_lens, n, max_len < class_max[c] ? max_len : class_max[c], encoding, bs_r1,
                   bs_r1, n, cnt + c, off + c, s_idx};
        unsigned int *cc = nullptr;
        unsigned g = (unsigned)blocks;
        if (ctr) {
            cc = ctr;
            zero_u32_kernel<<<1, 1, 0, st>>>(cc);
            MA_LAUNCHED();
        }
        MA_TRY(dispatch_map(c, p->max_indel > 0, lds, g, st, v, kn, ic, mo, fo, ro, cc, (int)claim));
    }
    return MA_OK;
}

// ---- the forward host loop (`-device cpu`)

// Pass 1 sizes the vote tables to the batch's largest count of voting
// postings of one read; pass 2 maps with one table per wave, the tables of all
// waves within this many bytes (fewer waves for batches holding repeat reads;
// one block at least).
constexpr int64_t FWD_TABLE_BUDGET = 4ll << 30;
constexpr int64_t FWD_ENTRY_BYTES = 16;  // tag u32 + count u32 + key u64

int fwd_launch(bool lds, unsigned g, hipStream_t st, const IndexView &v, const FwdArgs &a,
               const RecOut &out) {
    if (lds) w8::lds::fwd_map_kernel<<<g, BLOCK, 0, st>>>(v, a, out);
    else w8::glb::fwd_map_kernel<<<g, BLOCK, 0, st>>>(v, a, out);
    MA_LAUNCHED();
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
    ws->chunks_per_wave = 0;  // the tuned value at each launch, until ma_workspace_tune sets one
    int rc = ws->counter.alloc(256);
    if (rc != MA_OK) {
        delete ws;
        return rc;
    }
    *out = ws;
    return MA_OK;
}

int ma_workspace_set_persistent(ma_workspace *ws, int32_t persistent) {
    MA_REQUIRE(ws, "ma_workspace_set_persistent: null workspace");
    ws->persistent = persistent ? 1 : 0;
    return MA_OK;
}

int ma_workspace_tune(ma_workspace *ws, int32_t persistent, int32_t chunks_per_wave) {
    MA_REQUIRE(ws && (persistent == 0 || persistent == 1) && chunks_per_wave >= 1 && chunks_per_wave <= 64,
               "ma_workspace_tune: bad arguments");
    ws->persistent = persistent;
    ws->chunks_per_wave = chunks_per_wave;
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
    MA_REQUIRE(n <= ws->max_reads || ws->max_reads == 0, "ma_map_batch_dev: n exceeds workspace");
    return map_launch(ix, (hipStream_t)stream, d_bases, stride, d_lens, n, max_len, encoding, bs_r1, bs_r2, n_r1,
                      p, d_merged, d_fw, d_rc, ws->persistent, ws->chunks_per_wave,
                      ws->counter.as<unsigned int>(), &ws->split, "ma_map_batch_dev");
}

int ma_map_batch_dev_noalloc(ma_index *ix, void *stream, const uint8_t *d_bases, int64_t stride,
                             const uint32_t *d_lens, int64_t n, int32_t max_len, int32_t encoding,
                             int32_t bs_r1, int32_t bs_r2, int64_t n_r1, const ma_params *p,
                             ma_recs *d_merged, ma_recs *d_fw, ma_recs *d_rc) {
    MA_REQUIRE(ix && d_merged, "ma_map_batch_dev_noalloc: null index/output");
    unsigned int *ctr = nullptr;
    MA_TRY(noalloc_scratch(ix->device, &ctr));
    return map_launch(ix, (hipStream_t)stream, d_bases, stride, d_lens, n, max_len, encoding, bs_r1, bs_r2, n_r1,
                      p, d_merged, d_fw, d_rc, 1, 0, ctr, nullptr, "ma_map_batch_dev_noalloc");
}

int ma_map_batch(ma_index *ix, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                 int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                 int64_t n_r1, const ma_params *p, ma_recs *merged, ma_recs *fw, ma_recs *rc) {
    MA_REQUIRE(ix && merged, "ma_map_batch: null index/output");
    MA_REQUIRE(n >= 0, "ma_map_batch: n < 0");
    MA_TRY(check_map_params(ix, p, max_len));
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
    ws.chunks_per_wave = 0;
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

int ma_map_forward_dev(ma_index *ix, ma_workspace *ws, void *stream, const uint8_t *d_bases, int64_t stride,
                       const uint32_t *d_lens, int64_t n, int32_t max_len, int32_t bs_r1, int32_t bs_r2,
                       int64_t n_r1, const ma_params *p, ma_recs *d_out) {
    MA_REQUIRE(ix && ws && p && d_out, "ma_map_forward_dev: null index/workspace/params/output");
    MA_REQUIRE(n >= 0 && n < (1ll << 31), "ma_map_forward_dev: n out of range");
    MA_REQUIRE(max_len >= 1 && max_len <= 256 && stride >= max_len,
               "ma_map_forward_dev: max_len must be 1..256 and stride >= max_len");
    MA_REQUIRE(p->k == ix->k, "ma_map_forward_dev: k differs from the index's");
    MA_REQUIRE(p->stride >= 1 && p->max_occ >= 1 && p->max_diff >= 0 && p->max_soft >= 0,
               "ma_map_forward_dev: bad knobs");
    if (n == 0) return MA_OK;
    MA_REQUIRE(d_bases && d_lens, "ma_map_forward_dev: null input");
    MA_HIP(hipSetDevice(ix->device));
    hipStream_t st = (hipStream_t)stream;
    MA_TRY(ws->fwd_aux.reserve(256));
    unsigned int *ctr = ws->fwd_aux.as<unsigned int>();
    unsigned long long *mx = (unsigned long long *)((char *)ws->fwd_aux.p + 8);
    const IndexView v = index_view(ix);
    const RecOut out = rec_out(d_out);
    const bool lds = ix->n_contigs < COFF_LDS;
    FwdArgs a{};
    a.bases = d_bases;
    a.stride = stride;
    a.lens = d_lens;
    a.n = n;
    a.n_r1 = n_r1;
    a.max_len = max_len;
    a.bs_r1 = bs_r1;
    a.bs_r2 = bs_r2;
    a.k = p->k;
    a.seed_stride = p->stride;
    a.max_occ = p->max_occ;
    a.max_diff = p->max_diff;
    a.max_soft = p->max_soft;
    a.counter = ctr;
    a.claim = 16;
    a.max_total = mx;
    // pass 1: the largest count of voting postings of one read
    a.count_only = 1;
    MA_HIP(hipMemsetAsync(ws->fwd_aux.p, 0, 16, st));
    MA_TRY(fwd_launch(lds, (unsigned)(ix->n_cu * 4), st, v, a, out));
    unsigned long long max_total = 0;
    MA_HIP(hipMemcpyAsync(&max_total, mx, 8, hipMemcpyDeviceToHost, st));
    MA_HIP(hipStreamSynchronize(st));
    int cap_log2 = 6;
    while ((1ull << cap_log2) < 2 * max_total) cap_log2++;
    const int64_t per = 1ll << cap_log2;
    int64_t waves = FWD_TABLE_BUDGET / (per * FWD_ENTRY_BYTES) / WPB * WPB;
    const int64_t most = (int64_t)ix->n_cu * RESIDENT_WAVES_PER_CU;
    if (waves > most) waves = most;
    if (waves < WPB) waves = WPB;
    // tables: tags, counts and keys as three arrays of the buffer's entry
    // capacity; a wave's table is its slice. Tags only grow between clears
    // (tag 0 = never written), so an entry left by an earlier read or launch
    // never carries the tag of the read voting now.
    const int64_t need = waves * per;
    if (ws->fwd_entries < need) {
        MA_TRY(ws->fwd_tab.alloc((size_t)(need * FWD_ENTRY_BYTES)));
        ws->fwd_entries = need;
        ws->fwd_tag = 0;
        MA_HIP(hipMemsetAsync(ws->fwd_tab.p, 0, (size_t)need * 4, st));
    } else if (ws->fwd_tag > 0xFFFFFFFFull - (uint64_t)n - 1) {
        ws->fwd_tag = 0;
        MA_HIP(hipMemsetAsync(ws->fwd_tab.p, 0, (size_t)ws->fwd_entries * 4, st));
    }
    char *tb = (char *)ws->fwd_tab.p;
    a.tags = (uint32_t *)tb;
    a.cnts = (uint32_t *)(tb + ws->fwd_entries * 4);
    a.keys = (uint64_t *)(tb + ws->fwd_entries * 8);
    a.cap_log2 = cap_log2;
    a.tag_base = (uint32_t)ws->fwd_tag;
    a.count_only = 0;
    MA_HIP(hipMemsetAsync(ws->fwd_aux.p, 0, 4, st));
    MA_TRY(fwd_launch(lds, (unsigned)(waves / WPB), st, v, a, out));
    ws->fwd_tag += (uint64_t)n + 1;
    return MA_OK;
}

int ma_map_forward(ma_index *ix, const uint8_t *bases, int64_t stride, const uint32_t *lens, int64_t n,
                   int32_t max_len, int32_t bs_r1, int32_t bs_r2, int64_t n_r1, const ma_params *p, ma_recs *out) {
    MA_REQUIRE(ix && p && out, "ma_map_forward: null index/params/output");
    MA_REQUIRE(n >= 0, "ma_map_forward: n < 0");
    if (n == 0) return MA_OK;
    MA_REQUIRE(bases && lens, "ma_map_forward: null input");
    for (int64_t i = 0; i < n; i++)
        MA_REQUIRE((int64_t)lens[i] <= max_len, "ma_map_forward: a read is longer than max_len");
    MA_HIP(hipSetDevice(ix->device));
    hipStream_t st = ix->stream;
    DevBuf d_in, d_lens, d_out;
    MA_TRY(d_in.alloc((size_t)(n * stride)));
    MA_TRY(d_lens.alloc((size_t)n * 4));
    MA_TRY(d_out.alloc((size_t)n * 4 * 10));
    char *b = (char *)d_out.p;
    ma_recs dr;
    dr.cid = (int32_t *)(b + 0 * n * 4);
    dr.pos = (uint32_t *)(b + 1 * n * 4);
    dr.mapq = (uint32_t *)(b + 2 * n * 4);
    dr.flag = (int32_t *)(b + 3 * n * 4);
    dr.sl = (uint32_t *)(b + 4 * n * 4);
    dr.sr = (uint32_t *)(b + 5 * n * 4);
    dr.nm = (uint32_t *)(b + 6 * n * 4);
    dr.iop = (uint32_t *)(b + 7 * n * 4);
    dr.iat = (uint32_t *)(b + 8 * n * 4);
    dr.ilen = (uint32_t *)(b + 9 * n * 4);
    MA_HIP(hipMemcpyAsync(d_in.p, bases, (size_t)(n * stride), hipMemcpyHostToDevice, st));
    MA_HIP(hipMemcpyAsync(d_lens.p, lens, (size_t)n * 4, hipMemcpyHostToDevice, st));
    ma_workspace ws;
    ws.ix = ix;
    MA_TRY(ma_map_forward_dev(ix, &ws, st, d_in.as<uint8_t>(), stride, d_lens.as<uint32_t>(), n, max_len, bs_r1,
                              bs_r2, n_r1, p, &dr));
    void *dst[10] = {out->cid, out->pos, out->mapq, out->flag, out->sl, out->sr, out->nm, out->iop, out->iat, out->ilen};
    for (int f = 0; f < 10; f++)
        if (dst[f]) MA_HIP(hipMemcpyAsync(dst[f], b + (size_t)f * n * 4, (size_t)n * 4, hipMemcpyDeviceToHost, st));
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
