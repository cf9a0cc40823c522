// Library-level entry points: errors, device info, memory, streams, events.
#include <string.h>

#include <atomic>
#include <string>

#include "ma_common.hpp"

namespace ma {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

// Launch knobs the autotuner sets (process-wide; every value gives the same
// results). Defaults are the hand-set values the kernels were first tuned at.
namespace {
struct TuneSpec {
    int64_t def, lo, hi;
};
constexpr TuneSpec TUNE_SPEC[MA_TUNE_COUNT] = {
    {16, 8, 16},  // MA_TUNE_SORT_TILE_ITEMS: keys per thread of a radix tile (8 or 16)
    {16, 1, 64},  // MA_TUNE_MAP_MAX_CLAIM: most reads a wave claims per counter atomic
    {8, 1, 16},   // MA_TUNE_MAP_BLOCKS_PER_CU: persistent map grid, blocks per CU
    {1, 1, 64},   // MA_TUNE_MAP_CHUNKS_PER_WAVE: default read chunks per resident wave
    {0, 0, 1},    // MA_TUNE_MAP_CLASS_SPLIT: map uneven batches per read-length class
};
std::atomic<int64_t> g_tune[MA_TUNE_COUNT] = {{16}, {16}, {8}, {1}, {0}};
}  // namespace

int64_t tune_get(int knob) { return g_tune[knob].load(std::memory_order_relaxed); }
}  // namespace ma

using namespace ma;

extern "C" {

const char *ma_last_error(void) { return g_last_error.c_str(); }

const char *ma_version(void) { return "mialign 0.1 gfx950"; }

int ma_tune_set(int32_t knob, int64_t value) {
    MA_REQUIRE(knob >= 0 && knob < MA_TUNE_COUNT, "ma_tune_set: unknown knob");
    const TuneSpec &t = TUNE_SPEC[knob];
    MA_REQUIRE(value >= t.lo && value <= t.hi, "ma_tune_set: value out of range");
    MA_REQUIRE(knob != MA_TUNE_SORT_TILE_ITEMS || value == 8 || value == 16, "ma_tune_set: sort tile items are 8 or 16");
    g_tune[knob].store(value, std::memory_order_relaxed);
    return MA_OK;
}

int ma_tune_get(int32_t knob, int64_t *value, int64_t *default_value) {
    MA_REQUIRE(knob >= 0 && knob < MA_TUNE_COUNT, "ma_tune_get: unknown knob");
    if (value) *value = g_tune[knob].load(std::memory_order_relaxed);
    if (default_value) *default_value = TUNE_SPEC[knob].def;
    return MA_OK;
}

void ma_params_default(ma_params *p) {
    p->k = 21;
    p->stride = 5;
    p->max_occ = 16384;
    p->vote_occ = 256;
    p->vote_fast = 256;
    p->max_diff = 6;
    p->max_soft = 8;
    p->max_indel = 0;
    p->do_rc = 1;
}

int ma_device_count(int *n) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *n = c;
    return MA_OK;
}

int ma_device_info(int device, char *name, int name_len, int64_t *hbm_total, int64_t *hbm_free,
                   int32_t *n_cu, int32_t *lds_per_block) {
    hipDeviceProp_t prop;
    MA_HIP(hipGetDeviceProperties(&prop, device));
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    MA_HIP(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    MA_HIP(hipMemGetInfo(&fr, &tot));
    if (hbm_total) *hbm_total = (int64_t)tot;
    if (hbm_free) *hbm_free = (int64_t)fr;
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (lds_per_block) *lds_per_block = (int32_t)prop.sharedMemPerBlock;
    return MA_OK;
}

int ma_dev_alloc(int device, int64_t bytes, void **ptr) {
    MA_REQUIRE(bytes >= 0 && ptr, "ma_dev_alloc: bad arguments");
    MA_HIP(hipSetDevice(device));
    MA_HIP(hipMalloc(ptr, (size_t)(bytes > 0 ? bytes : 256)));
    return MA_OK;
}

int ma_dev_free(void *ptr) {
    if (ptr) MA_HIP(hipFree(ptr));
    return MA_OK;
}

int ma_memcpy_h2d(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes <= 0) return MA_OK;
    MA_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    if (!stream) MA_HIP(hipStreamSynchronize(nullptr));
    return MA_OK;
}

int ma_memcpy_d2h(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes <= 0) return MA_OK;
    MA_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    if (!stream) MA_HIP(hipStreamSynchronize(nullptr));
    return MA_OK;
}

int ma_memcpy_d2d(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes <= 0) return MA_OK;
    MA_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return MA_OK;
}

int ma_stream_create(int device, void **stream) {
    MA_HIP(hipSetDevice(device));
    hipStream_t s;
    MA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return MA_OK;
}

int ma_stream_destroy(void *stream) {
    if (stream) MA_HIP(hipStreamDestroy((hipStream_t)stream));
    return MA_OK;
}

int ma_stream_sync(void *stream) {
    MA_HIP(hipStreamSynchronize((hipStream_t)stream));
    return MA_OK;
}

int ma_event_create(void **ev) {
    hipEvent_t e;
    MA_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return MA_OK;
}

int ma_event_destroy(void *ev) {
    if (ev) MA_HIP(hipEventDestroy((hipEvent_t)ev));
    return MA_OK;
}

int ma_event_record(void *ev, void *stream) {
    MA_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return MA_OK;
}

int ma_event_sync(void *ev) {
    MA_HIP(hipEventSynchronize((hipEvent_t)ev));
    return MA_OK;
}

int ma_event_elapsed_ms(void *start, void *stop, float *ms) {
    MA_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return MA_OK;
}

int ma_graph_begin(void *stream) {
    MA_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return MA_OK;
}

int ma_graph_end(void *stream, void **graph_exec, int64_t *n_nodes) {
    MA_REQUIRE(graph_exec, "ma_graph_end: null output");
    hipGraph_t g = nullptr;
    MA_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    size_t nn = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &nn);
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_error(std::string("ma_graph_end: ") + hipGetErrorString(e));
        return MA_ERR_HIP;
    }
    if (n_nodes) *n_nodes = (int64_t)nn;
    *graph_exec = (void *)x;
    return MA_OK;
}

int ma_graph_launch(void *graph_exec, void *stream) {
    MA_REQUIRE(graph_exec, "ma_graph_launch: null graph");
    MA_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return MA_OK;
}

int ma_graph_destroy(void *graph_exec) {
    if (graph_exec) MA_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return MA_OK;
}

int ma_event_query(void *ev) {
    hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    set_error(std::string("hipEventQuery: ") + hipGetErrorString(e));
    return MA_ERR_HIP;
}

}  // extern "C"
