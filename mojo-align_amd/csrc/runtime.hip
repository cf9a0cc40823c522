// Library-level entry points: errors, device info, memory, streams, events.
#include <string.h>

#include <string>

#include "ma_common.hpp"

namespace ma {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace ma

using namespace ma;

extern "C" {

const char *ma_last_error(void) { return g_last_error.c_str(); }

const char *ma_version(void) { return "mialign 0.1 gfx950"; }

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

int ma_event_query(void *ev) {
    hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    set_error(std::string("hipEventQuery: ") + hipGetErrorString(e));
    return MA_ERR_HIP;
}

}  // extern "C"
