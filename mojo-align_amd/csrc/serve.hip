// Served request loop: continuous batching over a resident index.
//
// Requests (read batches) are queued on submit and cut into pieces of at most
// `launch_reads` reads (split also where a paired request changes conversion,
// so every piece has one conversion). Whenever one of the server's `slots`
// streams is free, a launch is assembled from the waiting pieces in deadline
// order — deadline = arrival + w x the request's estimated stand-alone
// service time (w = 0 is arrival order; w > 0 lets a small request pass a
// large one that arrived shortly before it, without starving the large one) —
// merging compatible pieces (same row layout, conversion and knobs) up to
// `launch_reads` reads. A launch copies each piece's rows into its region of
// the slot's device buffer, maps all the reads with one kernel, and copies
// each piece's records back into its request's host arrays (one 2D copy per
// piece when the request's record arrays sit at a constant pitch in one
// block). Launches on different slots overlap on the device. A read's result
// does not depend on the batch it is mapped in, so every request gets the same
// records alone or merged with others.
//
// Device memory. Everything the server holds counts against its budget: slot
// buffers (reserved at open, grown only on a free slot), the slots' map
// workspaces, the target table and a pool of arenas for sorted requests. Each
// time the server takes more device memory it asks the driver for the free
// HBM of that moment (hipMemGetInfo) and takes it only if a headroom stays
// free — admission by live free memory, as gpu-common/python/gpu_mem.py:
// 203-377 does before a resident load. Steady-state launches take nothing.
//
// Sorted requests (one sample's batch: the reference marks duplicates per
// sample, never across a device's mix of samples). A request submitted with
// keys gets an arena from the pool at submit (or is not admitted: the caller
// waits for completions and submits again, as gpu_mem.py:332-377 waits for
// free HBM). Each launch holding one of its pieces copies the piece's records
// into the arena on the device; the launch holding its last piece waits for
// the request's other launches (hipStreamWaitEvent), then builds the sort
// keys over the whole request (SE, or PE pairs), sorts and marks duplicates
// on its own stream, and copies is_dup and perm back. The arena returns to
// the pool when the request completes.
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ma_index.hpp"

using namespace ma;

namespace {

constexpr int64_t DEFAULT_LAUNCH_READS = 128 << 10;
constexpr int64_t RESERVE_READS_MAX = 1 << 20;  // larger launches grow the slot buffer on first use
// Stand-alone service-time model used for deadlines: fixed + per-read cost.
constexpr double EST_FIXED_MS = 0.25;
constexpr double EST_MS_PER_READ = 1.0e-5;  // 100M reads/s
constexpr double DEFAULT_SIZE_WEIGHT = 0.5;
constexpr int REC_FIELDS = 10;  // cid pos mapq flag sl sr nm iop iat ilen
constexpr int KEY_FIELDS = 7;   // cid .. nm: the fields the sort keys read
constexpr int64_t HBM_HEADROOM = 256ll << 20;  // free HBM the server leaves to everyone else
constexpr int64_t POOL_CHUNK = 64ll << 20;     // arena pool grows by chunks of at least this
constexpr int STATUS_SLOTS = 4096;             // sorted requests admitted at once

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

int64_t up256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// A sorted request's device memory: its records (seven rows at pitch n x 4)
// and lens gathered from its launches, the sort keys, the outputs and the
// sort's scratch.
struct ArenaLayout {
    size_t recs, lens, coord, dhi, dlo, score, pair, dup, perm, scratch, total;
    int64_t scratch_bytes;
};

ArenaLayout arena_layout(int64_t n) {
    ArenaLayout A{};
    size_t off = 0;
    auto take = [&](size_t b) {
        size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    const size_t nn = (size_t)n;
    A.recs = take(nn * 4 * KEY_FIELDS);
    A.lens = take(nn * 4);
    A.coord = take(nn * 8);
    A.dhi = take(nn * 8);
    A.dlo = take(nn * 8);
    A.score = take(nn * 4);
    A.pair = take(nn * 4);
    A.dup = take(nn * 4);
    A.perm = take(nn * 4);
    A.scratch_bytes = ma_sort_scratch_bytes(n);
    A.scratch = take((size_t)A.scratch_bytes);
    A.total = off;
    return A;
}

struct Arena {
    int chunk = -1;
    int64_t off = 0, bytes = 0;
    char *p = nullptr;
};

struct Req {
    int64_t ticket = 0;
    const uint8_t *bases = nullptr;
    int64_t stride = 0;
    const uint32_t *lens = nullptr;
    int64_t n = 0;
    int32_t max_len = 0, encoding = 0, bs_r1 = 0, bs_r2 = 0;
    int64_t n_r1 = 0;
    ma_params p{};
    ma_recs out{};
    int64_t out_pitch = 0;  // bytes between consecutive record arrays, 0 = not one block
    int64_t issued = 0;     // reads handed to launches so far
    int32_t running = 0;    // launches in flight that hold pieces of this request
    double t_submit = 0, deadline = 0, latency = 0;
    // sorted requests
    int32_t keys = 0;  // 0 map only, 1 SE keys, 2 PE keys
    int32_t do_markdup = 0;
    uint32_t *is_dup = nullptr, *perm = nullptr;  // host outputs
    Arena arena;
    int32_t status = -1;            // this request's two words in the pinned status block
    std::vector<int64_t> launches;  // launches that took pieces of it
};

struct Piece {
    Req *r;
    int64_t off, cnt;  // request reads [off, off + cnt)
    int64_t at;        // first row of the piece in the launch
};

struct Launch {
    int slot = -1;
    hipEvent_t done = nullptr;
    std::vector<Piece> pieces;
};

struct Chunk {
    std::unique_ptr<ma::DevBuf> buf;
    std::map<int64_t, int64_t> free_at;  // offset -> bytes of each free range
};

int64_t launch_bytes(int64_t n, int64_t stride) {
    return up256(n * stride) + up256(n * 4) + up256(n * 4 * REC_FIELDS);
}

// Conversion of the request's next piece, and how many reads it may cover.
int32_t piece_mode(const Req &r, int64_t *limit) {
    if (r.issued < r.n_r1) {
        *limit = std::min(r.n, r.n_r1) - r.issued;
        return r.bs_r1;
    }
    *limit = r.n - r.issued;
    return r.bs_r2;
}

bool compatible(const Req &a, int32_t mode_a, const Req &b, int32_t mode_b) {
    return a.stride == b.stride && a.max_len == b.max_len && a.encoding == b.encoding && mode_a == mode_b &&
           memcmp(&a.p, &b.p, sizeof(ma_params)) == 0;
}

int64_t record_pitch(const ma_recs &o, int64_t n) {
    const char *f[REC_FIELDS] = {(const char *)o.cid, (const char *)o.pos, (const char *)o.mapq,
                                 (const char *)o.flag, (const char *)o.sl, (const char *)o.sr,
                                 (const char *)o.nm, (const char *)o.iop, (const char *)o.iat,
                                 (const char *)o.ilen};
    for (int i = 0; i < REC_FIELDS; i++)
        if (!f[i]) return 0;
    const int64_t pitch = f[1] - f[0];
    if (pitch < n * 4) return 0;
    for (int i = 2; i < REC_FIELDS; i++)
        if (f[i] - f[i - 1] != pitch) return 0;
    return pitch;
}

}  // namespace

struct ma_server {
    ma_index *ix = nullptr;
    int32_t max_len = 0;
    int64_t budget = 0;    // cap on the device bytes the server holds (0: free HBM only)
    int64_t free_hbm = 0;  // the driver's free-HBM figure at the last admission
    int64_t launch_reads = DEFAULT_LAUNCH_READS;
    double size_weight = DEFAULT_SIZE_WEIGHT;  // deadline = arrival + size_weight x estimate
    std::vector<hipStream_t> streams;
    std::vector<ma_workspace *> ws;
    std::vector<std::unique_ptr<ma::DevBuf>> slot_buf;  // grow-only per-slot device memory
    std::vector<int64_t> slot_launch;                   // -1 = free
    std::vector<Chunk> pool;                            // sorted requests' arenas
    ma::DevBuf tid;                                     // int32 target of each index contig
    int32_t n_tid = 0;
    uint32_t *status_host = nullptr;  // pinned, two sort error words per sorted request
    std::vector<int32_t> status_free;
    std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;  // queued or running
    std::vector<Req *> waiting;                              // requests with reads not yet issued
    std::deque<int64_t> completed;                           // finished tickets not yet reported
    std::unordered_map<int64_t, double> done_latency;        // finished ticket -> ms
    std::unordered_map<int64_t, Launch> launches;
    int64_t next_ticket = 0;
    int64_t next_launch = 0;
    std::mutex mu;
};

namespace {

int64_t held(const ma_server *s) {
    int64_t b = (int64_t)s->tid.bytes;
    for (auto &x : s->slot_buf) b += (int64_t)x->bytes;
    for (auto w : s->ws) b += (int64_t)(w->counter.bytes + w->split.bytes);
    for (auto &c : s->pool) b += (int64_t)c.buf->bytes;
    return b;
}

// Give back pool chunks no arena uses.
void release_idle_chunks(ma_server *s) {
    for (auto &c : s->pool)
        if (c.buf->p && c.free_at.size() == 1 && c.free_at.begin()->first == 0 &&
            c.free_at.begin()->second == (int64_t)c.buf->bytes) {
            c.buf->release();
            c.free_at.clear();
        }
}

// May the server take `extra` more device bytes now? Within the budget with
// what it holds, and within the driver's free HBM of this moment less the
// headroom. Idle pool chunks are given back first when either falls short.
int admit(ma_server *s, int64_t extra, bool *ok) {
    for (int attempt = 0; attempt < 2; attempt++) {
        *ok = false;
        if (attempt) release_idle_chunks(s);
        if (s->budget > 0 && held(s) + extra > s->budget) continue;
        size_t fr = 0, tot = 0;
        MA_HIP(hipMemGetInfo(&fr, &tot));
        s->free_hbm = (int64_t)fr;
        if ((int64_t)fr >= extra + HBM_HEADROOM) {
            *ok = true;
            return MA_OK;
        }
    }
    return MA_OK;
}

// Grow a buffer no launch is using to at least `need` bytes (to `want` when
// that is admitted too). *ok = false when neither is admitted.
int grow(ma_server *s, ma::DevBuf &b, int64_t need, int64_t want, bool *ok) {
    *ok = true;
    if (b.p && (int64_t)b.bytes >= need) return MA_OK;
    for (int64_t target : {std::max(need, want), need}) {
        MA_TRY(admit(s, target - (int64_t)b.bytes, ok));
        if (!*ok) continue;
        const int rc = b.reserve((size_t)target);
        if (rc == MA_OK) return MA_OK;
        if (rc != MA_ERR_NOMEM) return rc;
        *ok = false;  // another allocation took the memory first
    }
    return MA_OK;
}

// First fit in the pool's chunks, else a new chunk (admitted like any growth).
int pool_take(ma_server *s, int64_t bytes, Arena *a, bool *ok) {
    bytes = up256(bytes);
    *ok = true;
    for (size_t c = 0; c < s->pool.size(); c++) {
        auto &fm = s->pool[c].free_at;
        for (auto it = fm.begin(); it != fm.end(); ++it)
            if (it->second >= bytes) {
                a->chunk = (int)c;
                a->off = it->first;
                a->bytes = bytes;
                a->p = (char *)s->pool[c].buf->p + a->off;
                const int64_t rest = it->second - bytes, at = it->first + bytes;
                fm.erase(it);
                if (rest > 0) fm[at] = rest;
                return MA_OK;
            }
    }
    for (int64_t sz : {std::max(bytes, POOL_CHUNK), bytes}) {
        MA_TRY(admit(s, sz, ok));
        if (!*ok) continue;
        Chunk ch;
        ch.buf.reset(new ma::DevBuf());
        const int rc = ch.buf->alloc((size_t)sz);
        if (rc == MA_ERR_NOMEM) {
            *ok = false;
            continue;
        }
        MA_TRY(rc);
        if (sz > bytes) ch.free_at[bytes] = sz - bytes;
        size_t slot = 0;
        while (slot < s->pool.size() && s->pool[slot].buf->p) slot++;  // reuse a released chunk's place
        if (slot == s->pool.size()) s->pool.emplace_back();
        s->pool[slot] = std::move(ch);
        a->chunk = (int)slot;
        a->off = 0;
        a->bytes = bytes;
        a->p = (char *)s->pool[slot].buf->p;
        return MA_OK;
    }
    return MA_OK;
}

void pool_give(ma_server *s, Arena &a) {
    if (a.chunk < 0) return;
    auto &fm = s->pool[a.chunk].free_at;
    int64_t off = a.off, bytes = a.bytes;
    auto nx = fm.lower_bound(off);
    if (nx != fm.end() && off + bytes == nx->first) {
        bytes += nx->second;
        nx = fm.erase(nx);
    }
    if (nx != fm.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += bytes;
            a = Arena();
            return;
        }
    }
    fm[off] = bytes;
    a = Arena();
}

// Size every slot buffer (no launch in flight) for a launch of launch_reads
// reads (at most RESERVE_READS_MAX), the slots together within half of what
// the budget leaves beside the server's other holdings (the other half is for
// sorted requests' arenas); what is not admitted is left for later.
int reserve_slots(ma_server *s, int64_t launch_reads) {
    int64_t reserve = launch_bytes(std::min(launch_reads, RESERVE_READS_MAX), s->max_len);
    if (s->budget > 0) {
        int64_t other = held(s);
        for (auto &b : s->slot_buf) other -= (int64_t)b->bytes;
        reserve = std::min(reserve, ((s->budget - other) / (2 * (int64_t)s->slot_buf.size())) & ~(int64_t)255);
    }
    for (auto &b : s->slot_buf) {
        if ((int64_t)b->bytes > reserve) b->release();
        bool ok;
        if (reserve > 0) MA_TRY(grow(s, *b, reserve, reserve, &ok));
    }
    return MA_OK;
}

void release_request(ma_server *s, Req *r) {
    pool_give(s, r->arena);
    if (r->status >= 0) s->status_free.push_back(r->status);
    r->status = -1;
}

// Keys, sort and duplicate marking over a sorted request's gathered records,
// then is_dup / perm and the sort's two error words back to the host.
int enqueue_sort(ma_server *s, Req *r, hipStream_t st) {
    const ArenaLayout A = arena_layout(r->n);
    char *a = r->arena.p;
    const size_t row = (size_t)r->n * 4;
    ma_recs m{};
    m.cid = (int32_t *)(a + A.recs + 0 * row);
    m.pos = (uint32_t *)(a + A.recs + 1 * row);
    m.mapq = (uint32_t *)(a + A.recs + 2 * row);
    m.flag = (int32_t *)(a + A.recs + 3 * row);
    m.sl = (uint32_t *)(a + A.recs + 4 * row);
    m.sr = (uint32_t *)(a + A.recs + 5 * row);
    m.nm = (uint32_t *)(a + A.recs + 6 * row);
    const uint32_t *lens = (const uint32_t *)(a + A.lens);
    uint64_t *coord = (uint64_t *)(a + A.coord), *dhi = (uint64_t *)(a + A.dhi), *dlo = (uint64_t *)(a + A.dlo);
    uint32_t *score = (uint32_t *)(a + A.score), *pair = (uint32_t *)(a + A.pair);
    uint32_t *dup = (uint32_t *)(a + A.dup), *perm = (uint32_t *)(a + A.perm);
    const int32_t *tid = (const int32_t *)s->tid.p;
    if (r->keys == 1)
        MA_TRY(ma_sort_keys_se_dev(st, r->n, &m, lens, tid, s->n_tid, 0, coord, dhi, dlo, score, pair));
    else
        MA_TRY(ma_sort_keys_pe_dev(st, r->n / 2, &m, lens, tid, s->n_tid, 0, coord, dhi, dlo, score, pair));
    MA_TRY(ma_sort_markdup_dev(st, r->n, coord, dhi, dlo, score, pair, dup, perm, r->do_markdup, nullptr,
                               a + A.scratch, A.scratch_bytes));
    size_t w0 = 0, w1 = 0;
    sort_error_words(r->n, &w0, &w1);
    uint32_t *sw = s->status_host + 2 * r->status;
    MA_HIP(hipMemcpyAsync(sw, a + A.scratch + w0, 4, hipMemcpyDeviceToHost, st));
    MA_HIP(hipMemcpyAsync(sw + 1, a + A.scratch + w1, 4, hipMemcpyDeviceToHost, st));
    MA_HIP(hipMemcpyAsync(r->is_dup, dup, row, hipMemcpyDeviceToHost, st));
    MA_HIP(hipMemcpyAsync(r->perm, perm, row, hipMemcpyDeviceToHost, st));
    return MA_OK;
}

// Assemble and enqueue launches while a slot is free and waiting work fits.
int pump(ma_server *s) {
    while (!s->waiting.empty()) {
        int slot = -1;
        for (size_t i = 0; i < s->slot_launch.size(); i++)
            if (s->slot_launch[i] < 0) {
                slot = (int)i;
                break;
            }
        if (slot < 0) return MA_OK;
        if (s->size_weight > 0)
            std::stable_sort(s->waiting.begin(), s->waiting.end(),
                             [](const Req *a, const Req *b) { return a->deadline < b->deadline; });
        Req *head = s->waiting.front();
        int64_t lim0;
        const int32_t mode = piece_mode(*head, &lim0);
        const int64_t stride = head->stride;
        ma::DevBuf &buf = *s->slot_buf[slot];
        const int64_t want = launch_bytes(std::min(s->launch_reads, RESERVE_READS_MAX), stride);
        std::vector<Piece> pieces;
        int64_t total = 0;
        for (Req *q : s->waiting) {
            if (total >= s->launch_reads) break;
            int64_t lim;
            const int32_t m = piece_mode(*q, &lim);
            if (q != head && !compatible(*head, mode, *q, m)) continue;
            // a request that fits a launch of its own is not split over two:
            // every launch lasts at least as long as its slowest read
            if (total > 0 && lim > s->launch_reads - total && lim <= s->launch_reads) continue;
            int64_t cnt = std::min(lim, s->launch_reads - total);
            while (cnt > 0) {  // the slot's buffer must hold the launch
                bool ok;
                MA_TRY(grow(s, buf, launch_bytes(total + cnt, stride), want, &ok));
                if (ok) break;
                cnt >>= 1;
            }
            if (cnt <= 0) break;
            pieces.push_back(Piece{q, q->issued, cnt, total});
            total += cnt;
            if (cnt < lim) break;  // launch full (reads or memory)
        }
        if (total == 0) {
            if (!s->launches.empty()) return MA_OK;  // the most urgent work waits for memory
            set_error("serve: not enough free HBM for a launch of the most urgent request (slot buffer of " +
                      std::to_string(buf.bytes) + " bytes; the server holds " + std::to_string(held(s)) +
                      ", budget " + std::to_string(s->budget) + ", free " + std::to_string(s->free_hbm) + ")");
            return MA_ERR_NOMEM;
        }
        MA_HIP(hipSetDevice(s->ix->device));
        hipStream_t st = s->streams[slot];
        const int64_t lid = s->next_launch;
        Launch L;
        L.slot = slot;
        char *b = (char *)buf.p;
        uint8_t *d_bases = (uint8_t *)b;
        uint32_t *d_lens = (uint32_t *)(b + up256(total * stride));
        char *rec = b + up256(total * stride) + up256(total * 4);
        for (const Piece &pc : pieces) {
            MA_HIP(hipMemcpyAsync(d_bases + pc.at * stride, pc.r->bases + pc.off * pc.r->stride,
                                  (size_t)(pc.cnt * pc.r->stride), hipMemcpyHostToDevice, st));
            MA_HIP(hipMemcpyAsync(d_lens + pc.at, pc.r->lens + pc.off, (size_t)pc.cnt * 4,
                                  hipMemcpyHostToDevice, st));
        }
        ma_recs dr;
        dr.cid = (int32_t *)(rec + 0 * total * 4);
        dr.pos = (uint32_t *)(rec + 1 * total * 4);
        dr.mapq = (uint32_t *)(rec + 2 * total * 4);
        dr.flag = (int32_t *)(rec + 3 * total * 4);
        dr.sl = (uint32_t *)(rec + 4 * total * 4);
        dr.sr = (uint32_t *)(rec + 5 * total * 4);
        dr.nm = (uint32_t *)(rec + 6 * total * 4);
        dr.iop = (uint32_t *)(rec + 7 * total * 4);
        dr.iat = (uint32_t *)(rec + 8 * total * 4);
        dr.ilen = (uint32_t *)(rec + 9 * total * 4);
        MA_TRY(ma_map_batch_dev(s->ix, s->ws[slot], st, d_bases, stride, d_lens, total, head->max_len,
                                head->encoding, mode, mode, total, &head->p, &dr, nullptr, nullptr));
        // all ten fields go back, so a caller's iop / iat / ilen arrays hold the
        // launch's values (zero for gapless launches) whatever they held before
        for (const Piece &pc : pieces) {
            Req *r = pc.r;
            const ma_recs &o = r->out;
            if (r->out_pitch) {
                MA_HIP(hipMemcpy2DAsync((char *)o.cid + pc.off * 4, (size_t)r->out_pitch, rec + pc.at * 4,
                                        (size_t)total * 4, (size_t)pc.cnt * 4, (size_t)REC_FIELDS,
                                        hipMemcpyDeviceToHost, st));
            } else {
                void *dst[REC_FIELDS] = {o.cid, o.pos, o.mapq, o.flag, o.sl, o.sr, o.nm, o.iop, o.iat, o.ilen};
                for (int f = 0; f < REC_FIELDS; f++)
                    if (dst[f])
                        MA_HIP(hipMemcpyAsync((char *)dst[f] + pc.off * 4, rec + ((size_t)f * total + pc.at) * 4,
                                              (size_t)pc.cnt * 4, hipMemcpyDeviceToHost, st));
            }
            if (!r->keys) continue;
            // gather the piece's records and lens into the request's arena
            const ArenaLayout A = arena_layout(r->n);
            MA_HIP(hipMemcpy2DAsync(r->arena.p + A.recs + pc.off * 4, (size_t)r->n * 4, rec + pc.at * 4,
                                    (size_t)total * 4, (size_t)pc.cnt * 4, (size_t)KEY_FIELDS,
                                    hipMemcpyDeviceToDevice, st));
            MA_HIP(hipMemcpyAsync(r->arena.p + A.lens + pc.off * 4, d_lens + pc.at, (size_t)pc.cnt * 4,
                                  hipMemcpyDeviceToDevice, st));
        }
        // sorts last (a wait on another slot holds back everything after it):
        // for each request whose last piece is here, its earlier launches, then its sort
        for (const Piece &pc : pieces) {
            Req *r = pc.r;
            if (!r->keys || pc.off + pc.cnt < r->n) continue;
            for (int64_t o_lid : r->launches) {
                auto it = s->launches.find(o_lid);
                if (it != s->launches.end() && it->second.slot != slot)
                    MA_HIP(hipStreamWaitEvent(st, it->second.done, 0));
            }
            MA_TRY(enqueue_sort(s, r, st));
        }
        MA_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
        MA_HIP(hipEventRecord(L.done, st));
        for (const Piece &pc : pieces) {
            pc.r->issued += pc.cnt;
            pc.r->running++;
            pc.r->launches.push_back(lid);
        }
        s->waiting.erase(std::remove_if(s->waiting.begin(), s->waiting.end(),
                                        [](const Req *q) { return q->issued >= q->n; }),
                         s->waiting.end());
        L.pieces = std::move(pieces);
        s->next_launch++;
        s->slot_launch[slot] = lid;
        s->launches[lid] = std::move(L);
    }
    return MA_OK;
}

// Free a finished launch's slot; requests with nothing left complete (a
// sorted request also gives its arena back, after its sort's error words are
// checked).
int retire(ma_server *s, int64_t lid) {
    auto it = s->launches.find(lid);
    if (it == s->launches.end()) return MA_OK;
    Launch &L = it->second;
    MA_HIP(hipEventDestroy(L.done));
    s->slot_launch[L.slot] = -1;
    const double t = now_ms();
    int rc = MA_OK;
    for (const Piece &pc : L.pieces) {
        Req *r = pc.r;
        if (--r->running == 0 && r->issued >= r->n) {
            if (r->keys && (s->status_host[2 * r->status] || s->status_host[2 * r->status + 1])) {
                set_error("serve: request " + std::to_string(r->ticket) +
                          ": a radix pass of its sort timed out waiting for a preceding tile (is_dup / perm invalid)");
                rc = MA_ERR_HIP;
            }
            release_request(s, r);
            s->done_latency[r->ticket] = t - r->t_submit;
            s->completed.push_back(r->ticket);
            s->reqs.erase(r->ticket);
        }
    }
    s->launches.erase(it);
    return rc;
}

// Retire every launch whose event has fired.
int reap(ma_server *s) {
    std::vector<int64_t> ready;
    for (auto &kv : s->launches) {
        hipError_t e = hipEventQuery(kv.second.done);
        if (e == hipSuccess) ready.push_back(kv.first);
        else if (e != hipErrorNotReady) {
            set_error(std::string("serve: ") + hipGetErrorString(e));
            return MA_ERR_HIP;
        }
    }
    std::sort(ready.begin(), ready.end());
    int rc = MA_OK;
    for (int64_t lid : ready) {
        const int r = retire(s, lid);
        if (rc == MA_OK) rc = r;
    }
    return rc;
}

// Block until some launch finishes (the oldest in flight).
int wait_one(ma_server *s) {
    if (s->launches.empty()) return MA_OK;
    int64_t oldest = s->launches.begin()->first;
    for (auto &kv : s->launches) oldest = std::min(oldest, kv.first);
    MA_HIP(hipEventSynchronize(s->launches[oldest].done));
    return MA_OK;
}

void forget_completed(ma_server *s, int64_t ticket) {
    for (auto it = s->completed.begin(); it != s->completed.end(); ++it)
        if (*it == ticket) {
            s->completed.erase(it);
            break;
        }
}

int64_t submit(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens, int64_t n,
               int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2, int64_t n_r1, const ma_params *p,
               ma_recs *merged_out, int32_t keys, int32_t do_markdup, uint32_t *is_dup, uint32_t *perm) {
    MA_REQUIRE(s && bases && lens && merged_out && p && n > 0 && n_r1 >= 0 && stride > 0,
               "ma_serve_submit: bad arguments");
    MA_REQUIRE(max_len <= s->max_len, "ma_serve_submit: max_len above the server's");
    MA_REQUIRE(keys >= 0 && keys <= 2, "ma_serve_submit_sorted: keys must be 0, 1 (SE) or 2 (PE)");
    MA_REQUIRE(!keys || (is_dup && perm), "ma_serve_submit_sorted: null is_dup / perm");
    MA_REQUIRE(keys != 2 || (n % 2 == 0 && n_r1 == n / 2),
               "ma_serve_submit_sorted: PE keys need R1 rows [0, n/2) and R2 rows [n/2, n)");
    MA_REQUIRE(!keys || n <= 0xFFFFFFFFll, "ma_serve_submit_sorted: n exceeds u32 record ids");
    MA_TRY(check_map_params(s->ix, p, max_len));
    std::lock_guard<std::mutex> lk(s->mu);
    MA_HIP(hipSetDevice(s->ix->device));
    std::unique_ptr<Req> r(new Req());
    if (keys) {  // admission: the arena now, or not at all
        MA_REQUIRE(s->n_tid > 0, "ma_serve_submit_sorted: no target table");
        if (s->status_free.empty()) {
            set_error("ma_serve_submit_sorted: " + std::to_string(STATUS_SLOTS) + " sorted requests already in flight");
            return MA_ERR_NOMEM;
        }
        bool ok;
        const int64_t need = (int64_t)arena_layout(n).total;
        MA_TRY(pool_take(s, need, &r->arena, &ok));
        if (!ok) {
            set_error("ma_serve_submit_sorted: not admitted: its arena of " + std::to_string(need) +
                      " bytes does not fit (the server holds " + std::to_string(held(s)) + " bytes, budget " +
                      std::to_string(s->budget) + ", free HBM " + std::to_string(s->free_hbm) + ")");
            return MA_ERR_NOMEM;
        }
        r->status = s->status_free.back();
        s->status_free.pop_back();
        r->keys = keys;
        r->do_markdup = do_markdup ? 1 : 0;
        r->is_dup = is_dup;
        r->perm = perm;
    }
    r->ticket = s->next_ticket++;
    r->bases = bases;
    r->stride = stride;
    r->lens = lens;
    r->n = n;
    r->max_len = max_len;
    r->encoding = encoding;
    r->bs_r1 = bs_r1;
    r->bs_r2 = bs_r2;
    r->n_r1 = n_r1;
    r->p = *p;
    r->out = *merged_out;
    r->out_pitch = record_pitch(r->out, n);
    r->t_submit = now_ms();
    r->deadline = r->t_submit + s->size_weight * (EST_FIXED_MS + EST_MS_PER_READ * (double)n);
    const int64_t t = r->ticket;
    Req *rp = r.get();
    s->waiting.push_back(rp);
    s->reqs[t] = std::move(r);
    int rc = reap(s);
    if (rc == MA_OK) rc = pump(s);
    if (rc != MA_OK && rp->issued == 0) {  // not started: the caller owns its arrays again
        s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), rp), s->waiting.end());
        release_request(s, rp);
        s->reqs.erase(t);
    }
    return rc != MA_OK ? rc : t;
}

}  // namespace

extern "C" {

int ma_serve_open(ma_index *ix, int32_t slots, int64_t hbm_budget, int32_t max_len,
                  ma_server **out) {
    MA_REQUIRE(ix && out && slots >= 1 && slots <= 64 && hbm_budget >= 0 && max_len >= 1 && max_len <= 256,
               "ma_serve_open: bad arguments");
    MA_HIP(hipSetDevice(ix->device));
    ma_server *s = new ma_server();
    s->ix = ix;
    s->max_len = max_len;
    s->budget = hbm_budget;
    int rc = ma_host_alloc((int64_t)STATUS_SLOTS * 2 * 4, (void **)&s->status_host);
    if (rc != MA_OK) {
        delete s;
        return rc;
    }
    for (int i = STATUS_SLOTS - 1; i >= 0; i--) s->status_free.push_back(i);
    // Targets: index contig c is target c until ma_serve_set_targets says otherwise.
    rc = s->tid.alloc((size_t)std::max<int64_t>(ix->n_contigs, 1) * 4);
    if (rc == MA_OK) {
        std::vector<int32_t> ident((size_t)ix->n_contigs);
        for (int64_t c = 0; c < ix->n_contigs; c++) ident[(size_t)c] = (int32_t)c;
        if (ix->n_contigs && hipMemcpy(s->tid.p, ident.data(), ident.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("ma_serve_open: target table copy failed");
            rc = MA_ERR_HIP;
        }
        s->n_tid = (int32_t)ix->n_contigs;
    }
    if (rc != MA_OK) {
        ma_serve_close(s);
        return rc;
    }
    for (int i = 0; i < slots; i++) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            ma_serve_close(s);
            set_error("ma_serve_open: stream creation failed");
            return MA_ERR_HIP;
        }
        s->streams.push_back(st);
        s->slot_launch.push_back(-1);
        ma_workspace *w = nullptr;
        rc = ma_workspace_create(ix, 0, max_len, &w);
        if (rc != MA_OK) {
            ma_serve_close(s);
            return rc;
        }
        w->persistent = 0;  // launches sized to their batch interleave on the device
        w->chunks_per_wave = 4;
        s->ws.push_back(w);
        s->slot_buf.emplace_back(new ma::DevBuf());
    }
    // Reserve each slot's launch buffer now (largest row layout: one byte per
    // base), within its share of the budget and the free HBM; a slot left
    // short grows on the first launch that needs it (or its launches shrink).
    rc = reserve_slots(s, s->launch_reads);
    if (rc != MA_OK) {
        ma_serve_close(s);
        return rc;
    }
    *out = s;
    return MA_OK;
}

int ma_serve_tune(ma_server *s, int64_t launch_reads, double size_weight, int32_t persistent,
                  int32_t chunks_per_wave) {
    MA_REQUIRE(s && launch_reads >= 1024 && size_weight >= 0 && size_weight <= 1000,
               "ma_serve_tune: bad arguments");
    std::lock_guard<std::mutex> lk(s->mu);
    MA_REQUIRE(s->launches.empty(), "ma_serve_tune: launches in flight");
    for (auto w : s->ws) MA_TRY(ma_workspace_tune(w, persistent, chunks_per_wave));
    s->launch_reads = launch_reads;
    s->size_weight = size_weight;
    MA_HIP(hipSetDevice(s->ix->device));
    return reserve_slots(s, launch_reads);
}

int ma_serve_set_targets(ma_server *s, const int32_t *tid_of_cid, int32_t n_tid) {
    MA_REQUIRE(s && n_tid >= 0 && (tid_of_cid || n_tid == 0), "ma_serve_set_targets: bad arguments");
    std::lock_guard<std::mutex> lk(s->mu);
    MA_REQUIRE(s->launches.empty() && s->waiting.empty(), "ma_serve_set_targets: requests in flight");
    MA_HIP(hipSetDevice(s->ix->device));
    bool ok;
    MA_TRY(grow(s, s->tid, (int64_t)std::max(n_tid, 1) * 4, 0, &ok));
    MA_REQUIRE(ok, "ma_serve_set_targets: not enough free HBM for the table");
    if (n_tid) MA_HIP(hipMemcpy(s->tid.p, tid_of_cid, (size_t)n_tid * 4, hipMemcpyHostToDevice));
    s->n_tid = n_tid;
    return MA_OK;
}

int64_t ma_serve_submit(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                        int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                        int64_t n_r1, const ma_params *p, ma_recs *merged_out) {
    return submit(s, bases, stride, lens, n, max_len, encoding, bs_r1, bs_r2, n_r1, p, merged_out, 0, 0,
                  nullptr, nullptr);
}

int64_t ma_serve_submit_sorted(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                               int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                               int64_t n_r1, const ma_params *p, ma_recs *merged_out, int32_t keys,
                               int32_t do_markdup, uint32_t *is_dup, uint32_t *perm) {
    MA_REQUIRE(keys == 1 || keys == 2, "ma_serve_submit_sorted: keys must be 1 (SE) or 2 (PE)");
    return submit(s, bases, stride, lens, n, max_len, encoding, bs_r1, bs_r2, n_r1, p, merged_out, keys,
                  do_markdup, is_dup, perm);
}

int ma_serve_memory(ma_server *s, int64_t *held_bytes, int64_t *budget, int64_t *free_hbm) {
    MA_REQUIRE(s, "ma_serve_memory: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    MA_HIP(hipSetDevice(s->ix->device));
    size_t fr = 0, tot = 0;
    MA_HIP(hipMemGetInfo(&fr, &tot));
    if (held_bytes) *held_bytes = held(s);
    if (budget) *budget = s->budget;
    if (free_hbm) *free_hbm = (int64_t)fr;
    return MA_OK;
}

int ma_serve_poll(ma_server *s, int64_t ticket, int32_t wait, double *latency_ms) {
    MA_REQUIRE(s, "ma_serve_poll: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    MA_REQUIRE(s->reqs.count(ticket) || s->done_latency.count(ticket), "ma_serve_poll: unknown ticket");
    while (true) {
        MA_TRY(reap(s));
        MA_TRY(pump(s));
        if (s->done_latency.count(ticket)) break;
        if (!wait) return 0;
        MA_TRY(wait_one(s));
    }
    if (latency_ms) *latency_ms = s->done_latency[ticket];
    s->done_latency.erase(ticket);
    forget_completed(s, ticket);
    return 1;
}

int ma_serve_reap(ma_server *s, int64_t *tickets, int32_t cap, int32_t wait) {
    MA_REQUIRE(s && (tickets || cap == 0) && cap >= 0, "ma_serve_reap: bad arguments");
    std::lock_guard<std::mutex> lk(s->mu);
    while (true) {
        MA_TRY(reap(s));
        MA_TRY(pump(s));
        if (!s->completed.empty() || !wait || s->launches.empty()) break;
        MA_TRY(wait_one(s));
    }
    int32_t n = 0;
    while (n < cap && !s->completed.empty()) {
        int64_t t = s->completed.front();
        s->completed.pop_front();
        s->done_latency.erase(t);
        tickets[n++] = t;
    }
    return n;
}

int ma_serve_close(ma_server *s) {
    if (!s) return MA_OK;
    for (auto &kv : s->launches) (void)hipEventSynchronize(kv.second.done);
    while (!s->launches.empty()) (void)retire(s, s->launches.begin()->first);
    for (auto st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (auto w : s->ws) ma_workspace_destroy(w);
    if (s->status_host) (void)ma_host_free(s->status_host);
    delete s;
    return MA_OK;
}

int ma_host_alloc(int64_t bytes, void **ptr) {
    MA_REQUIRE(ptr && bytes >= 0, "ma_host_alloc: bad arguments");
    MA_HIP(hipHostMalloc(ptr, (size_t)(bytes > 0 ? bytes : 64), hipHostMallocDefault));
    return MA_OK;
}

int ma_host_free(void *ptr) {
    if (ptr) MA_HIP(hipHostFree(ptr));
    return MA_OK;
}

}  // extern "C"
