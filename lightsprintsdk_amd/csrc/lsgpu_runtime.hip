// lsgpu_runtime.hip -- error reporting, device selection and memory plumbing of the
// liblsgpu C ABI.
#include "lsgpu_internal.h"
#include <cstdarg>
#include <cstring>

namespace lsg {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap; va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...)
{
    va_list ap; va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what, int line)
{
    set_error("HIP error %d (%s) in %s at lsgpu line %d", (int)e, hipGetErrorString(e), what, line);
    return e == hipErrorOutOfMemory ? LSG_ERR_OOM : LSG_ERR_DEVICE;
}

static int g_ready = -1;

bool device_ready()
{
    if (g_ready >= 0) return g_ready == 1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) { g_ready = 0; return false; }
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) { g_ready = 0; return false; }
    g_ready = (strncmp(prop.gcnArchName, "gfx950", 6) == 0) ? 1 : 0;
    if (!g_ready) set_error("device %d is %s, liblsgpu is built for gfx950 only", dev, prop.gcnArchName);
    return g_ready == 1;
}

int stack_size_for_depth(uint32_t depth)
{
    if (depth + 1 <= 32) return 32;
    if (depth + 1 <= 64) return 64;
    return 0;
}

} // namespace lsg

using namespace lsg;

extern "C" {

const char* lsg_last_error(void) { return g_err; }
const char* lsg_version(void) { return "lsgpu 0.1 gfx950"; }

int lsg_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int good = 0;
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) == hipSuccess && strncmp(p.gcnArchName, "gfx950", 6) == 0) good++;
    }
    return good;
}

int lsg_set_device(int device)
{
    LSG_HIP(hipSetDevice(device));
    g_ready = -1;
    return device_ready() ? LSG_OK : LSG_ERR_NO_DEVICE;
}

int lsg_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(LSG_ERR_INVALID, "lsg_malloc: null ptr");
    if (!device_ready()) return fail(LSG_ERR_NO_DEVICE, "lsg_malloc: no gfx950 device");
    LSG_HIP(hipMalloc(ptr, bytes ? bytes : 16));
    return LSG_OK;
}
int lsg_free(void* ptr) { if (ptr) LSG_HIP(hipFree(ptr)); return LSG_OK; }
int lsg_memcpy_h2d(void* dst, const void* src, size_t bytes) { LSG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); return LSG_OK; }
int lsg_memcpy_d2h(void* dst, const void* src, size_t bytes) { LSG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); return LSG_OK; }
int lsg_memset(void* dst, int value, size_t bytes) { LSG_HIP(hipMemset(dst, value, bytes)); return LSG_OK; }
int lsg_synchronize(void) { LSG_HIP(hipDeviceSynchronize()); return LSG_OK; }
int lsg_stream_create(void** s) { LSG_HIP(hipStreamCreate((hipStream_t*)s)); return LSG_OK; }
int lsg_stream_destroy(void* s) { LSG_HIP(hipStreamDestroy((hipStream_t)s)); return LSG_OK; }
int lsg_stream_synchronize(void* s) { LSG_HIP(hipStreamSynchronize((hipStream_t)s)); return LSG_OK; }
int lsg_event_create(void** e) { LSG_HIP(hipEventCreate((hipEvent_t*)e)); return LSG_OK; }
int lsg_event_destroy(void* e) { LSG_HIP(hipEventDestroy((hipEvent_t)e)); return LSG_OK; }
int lsg_event_record(void* e, void* s) { LSG_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s)); return LSG_OK; }
int lsg_event_elapsed_ms(void* a, void* b, float* ms)
{
    LSG_HIP(hipEventSynchronize((hipEvent_t)b));
    LSG_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return LSG_OK;
}

} // extern "C"
