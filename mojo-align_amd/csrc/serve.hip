// Served request loop: continuous batching over a resident index.
//
// Requests (read batches) are queued on submit. Whenever one of the server's
// `slots` streams is free, every compatible request waiting at the head of the
// queue (same row layout, conversion and knobs; single-end, so one conversion
// applies to all its reads) is merged into one launch, as long as the merged
// device bytes fit the free-HBM budget (the admission-by-free-memory idea of
// gpu-common/python/gpu_mem.py:203-377, applied per launch). A launch copies
// each request's rows into its region of one device allocation, maps all the
// reads with one kernel, and copies each request's records back into that
// request's host arrays. Requests arriving while launches run are taken by the
// next free slot, so their work overlaps the running launches. Each slot keeps
// a grow-only device buffer, so steady-state launches allocate nothing. A read's result
// does not depend on the batch it is mapped in, so every request gets the same
// records alone or merged.
#include <string.h>

#include <deque>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ma_index.hpp"

using namespace ma;

namespace {

constexpr int64_t MAX_MERGE_READS = 512 << 10;  // reads per merged launch (keeps launches short)

struct Pending {
    int64_t ticket;
    const uint8_t *bases;
    int64_t stride;
    const uint32_t *lens;
    int64_t n;
    int32_t max_len, encoding, bs_r1, bs_r2;
    int64_t n_r1;
    ma_params p;
    ma_recs out;
};

struct Launch {
    int slot = -1;
    int64_t bytes = 0;
    hipEvent_t done = nullptr;
    std::vector<int64_t> tickets;
};

int64_t up256(int64_t b) { return (b + 255) & ~(int64_t)255; }

int64_t launch_bytes(int64_t n, int64_t stride) { return up256(n * stride) + up256(n * 4) + up256(n * 4 * 7); }

bool compatible(const Pending &a, const Pending &b) {
    return a.stride == b.stride && a.max_len == b.max_len && a.encoding == b.encoding &&
           a.bs_r1 == b.bs_r1 && a.n_r1 >= a.n && b.n_r1 >= b.n &&
           memcmp(&a.p, &b.p, sizeof(ma_params)) == 0;
}

}  // namespace

struct ma_server {
    ma_index *ix = nullptr;
    int32_t max_len = 0;
    int64_t budget = 0;
    int64_t in_use = 0;
    std::vector<hipStream_t> streams;
    std::vector<ma_workspace *> ws;
    std::vector<ma::DevBuf *> slot_buf;  // grow-only per-slot device memory
    std::vector<int64_t> slot_launch;  // -1 = free
    std::deque<int64_t> completed;  // finished tickets not yet reported
    std::deque<Pending> pending;
    std::unordered_map<int64_t, Launch> launches;
    std::unordered_map<int64_t, int64_t> ticket_launch;  // -1 = queued
    int64_t next_ticket = 0;
    int64_t next_launch = 0;
    std::mutex mu;
};

namespace {

// Start launches while a slot is free and the queue head fits the budget.
int pump(ma_server *s) {
    while (!s->pending.empty()) {
        int slot = -1;
        for (size_t i = 0; i < s->slot_launch.size(); i++)
            if (s->slot_launch[i] < 0) {
                slot = (int)i;
                break;
            }
        if (slot < 0) return MA_OK;
        const Pending &head = s->pending.front();
        int64_t total = 0;
        size_t take = 0;
        for (size_t i = 0; i < s->pending.size(); i++) {
            const Pending &q = s->pending[i];
            if (i > 0 && (!compatible(head, q) || total + q.n > MAX_MERGE_READS)) break;
            if (s->in_use + launch_bytes(total + q.n, head.stride) > s->budget) break;
            total += q.n;
            take = i + 1;
        }
        if (take == 0) return MA_OK;  // head waits for memory
        MA_HIP(hipSetDevice(s->ix->device));
        hipStream_t st = s->streams[slot];
        Launch L;
        L.slot = slot;
        L.bytes = launch_bytes(total, head.stride);
        MA_TRY(s->slot_buf[slot]->reserve((size_t)L.bytes));
        char *b = (char *)s->slot_buf[slot]->p;
        uint8_t *d_bases = (uint8_t *)b;
        uint32_t *d_lens = (uint32_t *)(b + up256(total * head.stride));
        char *rec = b + up256(total * head.stride) + up256(total * 4);
        int64_t off = 0;
        for (size_t i = 0; i < take; i++) {
            const Pending &q = s->pending[i];
            MA_HIP(hipMemcpyAsync(d_bases + off * q.stride, q.bases, (size_t)(q.n * q.stride),
                                  hipMemcpyHostToDevice, st));
            MA_HIP(hipMemcpyAsync(d_lens + off, q.lens, (size_t)q.n * 4, hipMemcpyHostToDevice, st));
            off += q.n;
        }
        ma_recs dr;
        dr.cid = (int32_t *)(rec + 0 * total * 4);
        dr.pos = (uint32_t *)(rec + 1 * total * 4);
        dr.mapq = (uint32_t *)(rec + 2 * total * 4);
        dr.flag = (int32_t *)(rec + 3 * total * 4);
        dr.sl = (uint32_t *)(rec + 4 * total * 4);
        dr.sr = (uint32_t *)(rec + 5 * total * 4);
        dr.nm = (uint32_t *)(rec + 6 * total * 4);
        dr.iop = dr.iat = dr.ilen = nullptr;
        const int64_t n_r1 = take == 1 ? head.n_r1 : total;
        int rc = ma_map_batch_dev(s->ix, s->ws[slot], st, d_bases, head.stride, d_lens, total,
                                  head.max_len, head.encoding, head.bs_r1, head.bs_r2, n_r1, &head.p,
                                  &dr, nullptr, nullptr);
        if (rc != MA_OK) return rc;
        off = 0;
        for (size_t i = 0; i < take; i++) {
            const Pending &q = s->pending[i];
            void *dst[7] = {q.out.cid, q.out.pos, q.out.mapq, q.out.flag, q.out.sl, q.out.sr, q.out.nm};
            for (int f = 0; f < 7; f++)
                if (dst[f])
                    MA_HIP(hipMemcpyAsync(dst[f], rec + ((size_t)f * total + off) * 4, (size_t)q.n * 4,
                                          hipMemcpyDeviceToHost, st));
            off += q.n;
        }
        MA_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
        MA_HIP(hipEventRecord(L.done, st));
        const int64_t lid = s->next_launch++;
        for (size_t i = 0; i < take; i++) {
            L.tickets.push_back(s->pending.front().ticket);
            s->ticket_launch[s->pending.front().ticket] = lid;
            s->pending.pop_front();
        }
        s->in_use += L.bytes;
        s->slot_launch[slot] = lid;
        s->launches[lid] = L;
    }
    return MA_OK;
}

// Free a finished launch's memory and slot; its tickets become complete.
int retire(ma_server *s, int64_t lid) {
    auto it = s->launches.find(lid);
    if (it == s->launches.end()) return MA_OK;
    Launch &L = it->second;
    MA_HIP(hipEventDestroy(L.done));
    s->in_use -= L.bytes;
    s->slot_launch[L.slot] = -1;
    for (int64_t t : L.tickets) {
        s->ticket_launch[t] = -2;  // complete
        s->completed.push_back(t);
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

}  // namespace

extern "C" {

int ma_serve_open(ma_index *ix, int32_t slots, int64_t hbm_budget, int32_t max_len,
                  ma_server **out) {
    MA_REQUIRE(ix && out && slots >= 1 && slots <= 64 && hbm_budget > 0,
               "ma_serve_open: bad arguments");
    MA_HIP(hipSetDevice(ix->device));
    ma_server *s = new ma_server();
    s->ix = ix;
    s->max_len = max_len;
    s->budget = hbm_budget;
    for (int i = 0; i < slots; i++) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            ma_serve_close(s);
            set_error("ma_serve_open: stream creation failed");
            return MA_ERR_HIP;
        }
        s->streams.push_back(st);
        ma_workspace *w = nullptr;
        int rc = ma_workspace_create(ix, 0, max_len, &w);
        if (rc != MA_OK) {
            ma_serve_close(s);
            return rc;
        }
        w->persistent = 0;  // launches sized to their batch interleave on the device
        s->ws.push_back(w);
        s->slot_buf.push_back(new ma::DevBuf());
        s->slot_launch.push_back(-1);
    }
    *out = s;
    return MA_OK;
}

int64_t ma_serve_submit(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                        int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                        int64_t n_r1, const ma_params *p, ma_recs *merged_out) {
    MA_REQUIRE(s && bases && lens && merged_out && p && n > 0, "ma_serve_submit: bad arguments");
    MA_REQUIRE(max_len <= s->max_len, "ma_serve_submit: max_len above the server's");
    MA_REQUIRE(launch_bytes(n, stride) <= s->budget, "ma_serve_submit: request larger than the HBM budget");
    std::lock_guard<std::mutex> lk(s->mu);
    Pending q{s->next_ticket++, bases, stride, lens, n, max_len, encoding, bs_r1, bs_r2, n_r1, *p,
              *merged_out};
    s->pending.push_back(q);
    s->ticket_launch[q.ticket] = -1;
    MA_TRY(reap(s));
    MA_TRY(pump(s));
    return q.ticket;
}

int ma_serve_poll(ma_server *s, int64_t ticket, int32_t wait, double *latency_ms) {
    MA_REQUIRE(s, "ma_serve_poll: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->ticket_launch.find(ticket);
    MA_REQUIRE(it != s->ticket_launch.end(), "ma_serve_poll: unknown ticket");
    while (true) {
        MA_TRY(reap(s));
        MA_TRY(pump(s));
        int64_t lid = s->ticket_launch[ticket];
        if (lid == -2) break;
        if (!wait) return 0;
        if (lid >= 0) {
            MA_HIP(hipEventSynchronize(s->launches[lid].done));
        } else if (!s->launches.empty()) {
            MA_HIP(hipEventSynchronize(s->launches.begin()->second.done));
        }
    }
    if (latency_ms) *latency_ms = 0.0;
    s->ticket_launch.erase(ticket);
    for (auto it2 = s->completed.begin(); it2 != s->completed.end(); ++it2)
        if (*it2 == ticket) {
            s->completed.erase(it2);
            break;
        }
    return 1;
}

int ma_serve_reap(ma_server *s, int64_t *tickets, int32_t cap, int32_t wait) {
    MA_REQUIRE(s && (tickets || cap == 0) && cap >= 0, "ma_serve_reap: bad arguments");
    std::lock_guard<std::mutex> lk(s->mu);
    while (true) {
        MA_TRY(reap(s));
        MA_TRY(pump(s));
        if (!s->completed.empty() || !wait || s->launches.empty()) break;
        MA_HIP(hipEventSynchronize(s->launches.begin()->second.done));
    }
    int32_t n = 0;
    while (n < cap && !s->completed.empty()) {
        int64_t t = s->completed.front();
        s->completed.pop_front();
        s->ticket_launch.erase(t);
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
    for (auto b : s->slot_buf) delete b;
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
