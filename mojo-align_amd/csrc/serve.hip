// Served request loop: continuous batching over a resident index.
//
// A server owns `slots` HIP streams. A request (one read batch) is admitted
// when a slot is free and its device bytes fit the free-HBM budget (the
// admission-by-free-memory idea of gpu-common/python/gpu_mem.py:203-377,
// applied per request instead of per run). Admitted requests run their H2D
// copy, map kernel and D2H copy on their slot's stream, so requests that
// arrive while others are on the GPU overlap with them. Per-read results do
// not depend on the batch a read is in, so a read maps the same alone or in
// the mix. Device buffers come from the stream-ordered allocator and return
// to it when the request completes.
#include <chrono>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ma_index.hpp"

using namespace ma;

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int slot = -1;
    void *dmem = nullptr;
    int64_t bytes = 0;
    hipEvent_t done = nullptr;
    double t_submit = 0.0;
    double latency = -1.0;
    bool finished = false;
};

}  // namespace

struct ma_server {
    ma_index *ix = nullptr;
    int32_t max_len = 0;
    int64_t budget = 0;
    int64_t in_use = 0;
    std::vector<hipStream_t> streams;
    std::vector<ma_workspace *> ws;
    std::vector<int64_t> slot_ticket;  // -1 = free
    std::unordered_map<int64_t, Request> reqs;
    int64_t next_ticket = 0;
    std::mutex mu;
};

static int64_t request_bytes(int64_t n, int64_t stride) {
    auto up = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
    return up(n * stride) + up(n * 4) + up(n * 4 * 7);
}

// Completes a request whose event has fired: frees its memory and slot.
static int retire(ma_server *s, int64_t ticket, Request &r) {
    if (r.finished) return MA_OK;
    MA_HIP(hipFreeAsync(r.dmem, s->streams[r.slot]));
    MA_HIP(hipEventDestroy(r.done));
    r.done = nullptr;
    s->in_use -= r.bytes;
    s->slot_ticket[r.slot] = -1;
    r.finished = true;
    if (r.latency < 0) r.latency = now_ms() - r.t_submit;
    (void)ticket;
    return MA_OK;
}

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
        s->ws.push_back(w);
        s->slot_ticket.push_back(-1);
    }
    *out = s;
    return MA_OK;
}

int64_t ma_serve_submit(ma_server *s, const uint8_t *bases, int64_t stride, const uint32_t *lens,
                        int64_t n, int32_t max_len, int32_t encoding, int32_t bs_r1, int32_t bs_r2,
                        int64_t n_r1, const ma_params *p, ma_recs *merged_out) {
    MA_REQUIRE(s && bases && lens && merged_out && n > 0, "ma_serve_submit: bad arguments");
    MA_REQUIRE(max_len <= s->max_len, "ma_serve_submit: max_len above the server's");
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t need = request_bytes(n, stride);
    MA_REQUIRE(need <= s->budget, "ma_serve_submit: request larger than the HBM budget");
    int slot = -1;
    for (size_t i = 0; i < s->slot_ticket.size(); i++)
        if (s->slot_ticket[i] < 0) {
            slot = (int)i;
            break;
        }
    if (slot < 0 || s->in_use + need > s->budget) {
        set_error("ma_serve_submit: not admitted (no free slot or HBM budget)");
        return MA_ERR_NOMEM;
    }
    MA_HIP(hipSetDevice(s->ix->device));
    hipStream_t st = s->streams[slot];
    Request r;
    r.slot = slot;
    r.bytes = need;
    r.t_submit = now_ms();
    MA_HIP(hipMallocAsync(&r.dmem, (size_t)need, st));
    char *b = (char *)r.dmem;
    auto up = [](int64_t x) { return (x + 255) & ~(int64_t)255; };
    uint8_t *d_bases = (uint8_t *)b;
    uint32_t *d_lens = (uint32_t *)(b + up(n * stride));
    char *rec = b + up(n * stride) + up(n * 4);
    ma_recs dr;
    dr.cid = (int32_t *)(rec + 0 * n * 4);
    dr.pos = (uint32_t *)(rec + 1 * n * 4);
    dr.mapq = (uint32_t *)(rec + 2 * n * 4);
    dr.flag = (int32_t *)(rec + 3 * n * 4);
    dr.sl = (uint32_t *)(rec + 4 * n * 4);
    dr.sr = (uint32_t *)(rec + 5 * n * 4);
    dr.nm = (uint32_t *)(rec + 6 * n * 4);
    dr.iop = dr.iat = dr.ilen = nullptr;
    MA_HIP(hipMemcpyAsync(d_bases, bases, (size_t)(n * stride), hipMemcpyHostToDevice, st));
    MA_HIP(hipMemcpyAsync(d_lens, lens, (size_t)n * 4, hipMemcpyHostToDevice, st));
    s->ws[slot]->max_reads = 0;
    int rc = ma_map_batch_dev(s->ix, s->ws[slot], st, d_bases, stride, d_lens, n, max_len, encoding,
                              bs_r1, bs_r2, n_r1, p, &dr, nullptr, nullptr);
    if (rc != MA_OK) {
        (void)hipFreeAsync(r.dmem, st);
        return rc;
    }
    void *dst[7] = {merged_out->cid, merged_out->pos, merged_out->mapq, merged_out->flag,
                    merged_out->sl,  merged_out->sr,  merged_out->nm};
    for (int f = 0; f < 7; f++)
        if (dst[f]) MA_HIP(hipMemcpyAsync(dst[f], rec + (size_t)f * n * 4, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    MA_HIP(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
    MA_HIP(hipEventRecord(r.done, st));
    int64_t ticket = s->next_ticket++;
    s->in_use += need;
    s->slot_ticket[slot] = ticket;
    s->reqs[ticket] = r;
    return ticket;
}

int ma_serve_poll(ma_server *s, int64_t ticket, int32_t wait, double *latency_ms) {
    MA_REQUIRE(s, "ma_serve_poll: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    MA_REQUIRE(it != s->reqs.end(), "ma_serve_poll: unknown ticket");
    Request &r = it->second;
    if (!r.finished) {
        if (wait) {
            MA_HIP(hipEventSynchronize(r.done));
        } else {
            hipError_t e = hipEventQuery(r.done);
            if (e == hipErrorNotReady) return 0;
            if (e != hipSuccess) {
                set_error(std::string("ma_serve_poll: ") + hipGetErrorString(e));
                return MA_ERR_HIP;
            }
        }
        MA_TRY(retire(s, ticket, r));
    }
    if (latency_ms) *latency_ms = r.latency;
    s->reqs.erase(it);
    return 1;
}

int ma_serve_close(ma_server *s) {
    if (!s) return MA_OK;
    for (auto &kv : s->reqs) {
        if (!kv.second.finished && kv.second.done) {
            (void)hipEventSynchronize(kv.second.done);
            (void)retire(s, kv.first, kv.second);
        }
    }
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
