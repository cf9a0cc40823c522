// Served request loop: continuous batching over a resident index.
//
// Requests (read batches) are queued on submit and cut into pieces of at most
// `launch_reads` reads (split also where a paired request changes conversion,
// so every piece has one conversion). Whenever one of the server's `slots`
// streams is free, a launch is assembled from the waiting pieces in deadline
// order — deadline = arrival + w x the request's estimated stand-alone
// service time (w = 0 is arrival order; w > 0 lets a small request pass a
// large one that arrived shortly before it, without starving the large one) —
// merging compatible pieces (same row layout, conversion and
// knobs) up to `launch_reads` reads while the launch's device bytes fit the
// free-HBM budget (the admission-by-free-memory idea of
// gpu-common/python/gpu_mem.py:203-377, applied per launch). A launch copies
// each piece's rows into its region of the slot's device buffer, maps all the
// reads with one kernel, and copies each piece's records back into its
// request's host arrays (one 2D copy per piece when the request's seven record
// arrays sit at a constant pitch in one block). Launches on different slots
// overlap on the device. Slot buffers are reserved at open and only grow, so
// steady-state launches allocate nothing. A read's result does not depend on
// the batch it is mapped in, so every request gets the same records alone or
// merged with others.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
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

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

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
};

struct Piece {
    Req *r;
    int64_t off, cnt;  // request reads [off, off + cnt)
    int64_t at;        // first row of the piece in the launch
};

struct Launch {
    int slot = -1;
    int64_t bytes = 0;
    hipEvent_t done = nullptr;
    std::vector<Piece> pieces;
};

int64_t up256(int64_t b) { return (b + 255) & ~(int64_t)255; }

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
    int64_t budget = 0;
    int64_t in_use = 0;
    int64_t launch_reads = DEFAULT_LAUNCH_READS;
    double size_weight = DEFAULT_SIZE_WEIGHT;  // deadline = arrival + size_weight x estimate
    std::vector<hipStream_t> streams;
    std::vector<ma_workspace *> ws;
    std::vector<std::unique_ptr<ma::DevBuf>> slot_buf;  // grow-only per-slot device memory
    std::vector<int64_t> slot_launch;                   // -1 = free
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
        std::vector<Piece> pieces;
        int64_t total = 0;
        for (Req *q : s->waiting) {
            if (total >= s->launch_reads) break;
            int64_t lim;
            const int32_t m = piece_mode(*q, &lim);
            if (q != head && !compatible(*head, mode, *q, m)) continue;
            int64_t cnt = std::min(lim, s->launch_reads - total);
            while (cnt > 0 && s->in_use + launch_bytes(total + cnt, head->stride) > s->budget) cnt >>= 1;
            if (cnt <= 0) break;
            pieces.push_back(Piece{q, q->issued, cnt, total});
            total += cnt;
            if (cnt < lim) break;  // launch full (reads or memory)
        }
        if (total == 0) return MA_OK;  // the most urgent work waits for memory
        MA_HIP(hipSetDevice(s->ix->device));
        hipStream_t st = s->streams[slot];
        Launch L;
        L.slot = slot;
        L.bytes = launch_bytes(total, head->stride);
        MA_TRY(s->slot_buf[slot]->reserve((size_t)L.bytes));
        char *b = (char *)s->slot_buf[slot]->p;
        uint8_t *d_bases = (uint8_t *)b;
        uint32_t *d_lens = (uint32_t *)(b + up256(total * head->stride));
        char *rec = b + up256(total * head->stride) + up256(total * 4);
        for (const Piece &pc : pieces) {
            MA_HIP(hipMemcpyAsync(d_bases + pc.at * head->stride, pc.r->bases + pc.off * pc.r->stride,
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
        int rc = ma_map_batch_dev(s->ix, s->ws[slot], st, d_bases, head->stride, d_lens, total, head->max_len,
                                  head->encoding, mode, mode, total, &head->p, &dr, nullptr, nullptr);
        if (rc != MA_OK) return rc;
        // all ten fields go back, so a caller's iop / iat / ilen arrays hold the
        // launch's values (zero for gapless launches) whatever they held before
        const int n_fields = REC_FIELDS;
        for (const Piece &pc : pieces) {
            const ma_recs &o = pc.r->out;
            if (pc.r->out_pitch) {
                MA_HIP(hipMemcpy2DAsync((char *)o.cid + pc.off * 4, (size_t)pc.r->out_pitch, rec + pc.at * 4,
                                        (size_t)total * 4, (size_t)pc.cnt * 4, (size_t)n_fields,
                                        hipMemcpyDeviceToHost, st));
            } else {
                void *dst[REC_FIELDS] = {o.cid, o.pos, o.mapq, o.flag, o.sl, o.sr, o.nm, o.iop, o.iat, o.ilen};
                for (int f = 0; f < n_fields; f++)
                    if (dst[f])
                        MA_HIP(hipMemcpyAsync((char *)dst[f] + pc.off * 4, rec + ((size_t)f * total + pc.at) * 4,
                                              (size_t)pc.cnt * 4, hipMemcpyDeviceToHost, st));
            }
        }
        MA_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
        MA_HIP(hipEventRecord(L.done, st));
        for (const Piece &pc : pieces) {
            pc.r->issued += pc.cnt;
            pc.r->running++;
        }
        s->waiting.erase(std::remove_if(s->waiting.begin(), s->waiting.end(),
                                        [](const Req *q) { return q->issued >= q->n; }),
                         s->waiting.end());
        L.pieces = std::move(pieces);
        const int64_t lid = s->next_launch++;
        s->in_use += L.bytes;
        s->slot_launch[slot] = lid;
        s->launches[lid] = std::move(L);
    }
    return MA_OK;
}

// Free a finished launch's budget and slot; requests with nothing left complete.
int retire(ma_server *s, int64_t lid) {
    auto it = s->launches.find(lid);
    if (it == s->launches.end()) return MA_OK;
    Launch &L = it->second;
    MA_HIP(hipEventDestroy(L.done));
    s->in_use -= L.bytes;
    s->slot_launch[L.slot] = -1;
    const double t = now_ms();
    for (const Piece &pc : L.pieces) {
        Req *r = pc.r;
        if (--r->running == 0 && r->issued >= r->n) {
            s->done_latency[r->ticket] = t - r->t_submit;
            s->completed.push_back(r->ticket);
            s->reqs.erase(r->ticket);
        }
    }
    s->launches.erase(it);
    return MA_OK;
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
    for (int64_t lid : ready) MA_TRY(retire(s, lid));
    return MA_OK;
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

}  // namespace

extern "C" {

int ma_serve_open(ma_index *ix, int32_t slots, int64_t hbm_budget, int32_t max_len,
                  ma_server **out) {
    MA_REQUIRE(ix && out && slots >= 1 && slots <= 64 && hbm_budget > 0 && max_len >= 1 && max_len <= 256,
               "ma_serve_open: bad arguments");
    MA_HIP(hipSetDevice(ix->device));
    ma_server *s = new ma_server();
    s->ix = ix;
    s->max_len = max_len;
    s->budget = hbm_budget;
    // Reserve each slot's launch buffer now (largest row layout: one byte per
    // base), within the budget's share of a slot.
    const int64_t reserve = std::min(launch_bytes(s->launch_reads, max_len), hbm_budget / slots);
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
        int rc = ma_workspace_create(ix, 0, max_len, &w);
        if (rc == MA_OK) {
            w->persistent = 0;  // launches sized to their batch interleave on the device
            w->chunks_per_wave = 4;
            s->ws.push_back(w);
            s->slot_buf.emplace_back(new ma::DevBuf());
            rc = s->slot_buf.back()->reserve((size_t)reserve);
        }
        if (rc != MA_OK) {
            ma_serve_close(s);
            return rc;
        }
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
    const int64_t reserve = std::min(launch_bytes(std::min(launch_reads, RESERVE_READS_MAX), s->max_len),
                                     s->budget / (int64_t)s->slot_buf.size());
    for (auto &b : s->slot_buf) MA_TRY(b->reserve((size_t)reserve));
    return MA_OK;
}

int64_t ma_serve_submit(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                        int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                        int64_t n_r1, const ma_params *p, ma_recs *merged_out) {
    MA_REQUIRE(s && bases && lens && merged_out && p && n > 0 && n_r1 >= 0 && stride > 0,
               "ma_serve_submit: bad arguments");
    MA_REQUIRE(max_len <= s->max_len, "ma_serve_submit: max_len above the server's");
    MA_REQUIRE(launch_bytes(1024, stride) <= s->budget, "ma_serve_submit: HBM budget too small for a launch");
    MA_TRY(check_map_params(s->ix, p, max_len));
    std::lock_guard<std::mutex> lk(s->mu);
    std::unique_ptr<Req> r(new Req());
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
    s->waiting.push_back(r.get());
    s->reqs[t] = std::move(r);
    MA_TRY(reap(s));
    MA_TRY(pump(s));
    return t;
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
