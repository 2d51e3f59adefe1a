// Process-wide device context, error reporting and grow-only buffers.
#include "vx_common.hpp"

#include <cstdlib>

namespace vx {

namespace {
thread_local std::string g_error;
}

int32_t fail(int32_t code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

Context &context() {
    static Context *ctx = new Context();  // lives until process exit
    return *ctx;
}

int32_t ensure_init() {
    Context &c = context();
    std::lock_guard<std::recursive_mutex> lock(c.mutex);
    if (c.initialised) {
        if (c.init_status != VX_OK) return fail(c.init_status, "%s", c.init_message.c_str());
        return VX_OK;
    }
    c.initialised = true;
    auto cache_failure = [&](const char *what, hipError_t e) {
        char buf[512];
        snprintf(buf, sizeof buf, "HIP device initialisation failed: %s: %s", what, hipGetErrorString(e));
        c.init_status = VX_ERR_INIT;
        c.init_message = buf;
        return fail(VX_ERR_INIT, "%s", buf);
    };
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return cache_failure("hipGetDeviceCount", e);
    if (count < 1) {
        c.init_status = VX_ERR_INIT;
        c.init_message = "HIP device initialisation failed: no HIP device";
        return fail(VX_ERR_INIT, "%s", c.init_message.c_str());
    }
    if ((e = hipSetDevice(0)) != hipSuccess) return cache_failure("hipSetDevice(0)", e);
    // Default-priority non-blocking stream.
    if ((e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking)) != hipSuccess)
        return cache_failure("hipStreamCreateWithFlags", e);
    c.pinned_bytes = 4096;
    if ((e = hipHostMalloc(&c.pinned, c.pinned_bytes, hipHostMallocDefault)) != hipSuccess)
        return cache_failure("hipHostMalloc", e);
    c.init_status = VX_OK;
    return VX_OK;
}

int32_t buffer(const char *name, size_t bytes, void **out) {
    Context &c = context();
    if (bytes < 256) bytes = 256;
    DeviceBuffer &b = c.buffers[name];
    if (b.ptr && b.bytes >= bytes) {
        *out = b.ptr;
        return VX_OK;
    }
    if (b.ptr) {
        VX_HIP(hipStreamSynchronize(c.stream));
        VX_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    VX_HIP(hipMalloc(&b.ptr, bytes));
    b.bytes = bytes;
    *out = b.ptr;
    return VX_OK;
}

}  // namespace vx

extern "C" const char *vx_last_error(void) { return vx::g_error.c_str(); }

extern "C" int32_t vx_init(void) {
    vx::clear_error();
    return vx::ensure_init();
}
