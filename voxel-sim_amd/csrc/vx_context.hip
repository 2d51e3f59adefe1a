// Process-wide device context, error reporting and grow-only buffers.
#include "vx_common.hpp"

#include <algorithm>
#include <cstdlib>
#include <vector>

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
    rows_scratch_setup();
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

namespace {
// STREAM copy: one 16-byte vector per thread and step, grid-stride.
__global__ __launch_bounds__(256) void copy_vec4(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

}  // namespace vx

extern "C" int32_t vx_copy_bandwidth(uint64_t bytes, uint32_t reps, double *out_us) {
    using namespace vx;
    clear_error();
    if (!out_us || reps == 0 || bytes == 0 || bytes % 16 || bytes > (uint64_t(1) << 36))
        return fail(VX_ERR_INVALID_ARGUMENT, "copy size must be a positive multiple of 16 bytes up to 64 GiB, reps > 0");
    if (int32_t c = ensure_init()) return c;
    Context &ctx = context();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex);
    void *src = nullptr, *dst = nullptr;
    VX_HIP(hipMalloc(&src, bytes));
    hipError_t e = hipMalloc(&dst, bytes);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return fail(VX_ERR_DEVICE, "hipMalloc failed: %s", hipGetErrorString(e));
    }
    const size_t n = bytes / 16;
    const uint32_t grid = uint32_t(std::min<size_t>((n + 255) / 256, size_t(1) << 20));
    std::vector<hipEvent_t> ev(2 * reps);
    std::vector<double> us(reps);
    int32_t status = VX_OK;
    for (auto &x : ev)
        if (hipEventCreate(&x) != hipSuccess) status = fail(VX_ERR_DEVICE, "hipEventCreate failed");
    if (status == VX_OK && hipMemsetAsync(src, 0, bytes, ctx.stream) != hipSuccess) status = fail(VX_ERR_DEVICE, "hipMemsetAsync failed");
    if (status == VX_OK) {
        hipLaunchKernelGGL(copy_vec4, dim3(grid), dim3(256), 0, ctx.stream, (const float4 *)src, (float4 *)dst, n);
        for (uint32_t r = 0; r < reps; ++r) {
            (void)hipEventRecord(ev[2 * r], ctx.stream);
            hipLaunchKernelGGL(copy_vec4, dim3(grid), dim3(256), 0, ctx.stream, (const float4 *)src, (float4 *)dst, n);
            (void)hipEventRecord(ev[2 * r + 1], ctx.stream);
        }
        if ((e = hipStreamSynchronize(ctx.stream)) != hipSuccess) status = fail(VX_ERR_DEVICE, "copy failed: %s", hipGetErrorString(e));
    }
    for (uint32_t r = 0; status == VX_OK && r < reps; ++r) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]) != hipSuccess) status = fail(VX_ERR_DEVICE, "hipEventElapsedTime failed");
        us[r] = 1e3 * double(ms);
    }
    for (auto &x : ev)
        if (x) (void)hipEventDestroy(x);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (status != VX_OK) return status;
    std::sort(us.begin(), us.end());
    *out_us = us[reps / 2];
    return VX_OK;
}

extern "C" const char *vx_last_error(void) { return vx::g_error.c_str(); }

extern "C" int32_t vx_set_launch_mode(int32_t mode) {
    using namespace vx;
    clear_error();
    if (mode != VX_LAUNCH_GRAPHS && mode != VX_LAUNCH_DIRECT)
        return fail(VX_ERR_INVALID_ARGUMENT, "launch mode must be VX_LAUNCH_GRAPHS or VX_LAUNCH_DIRECT");
    Context &ctx = context();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex);
    ctx.direct_launch = mode == VX_LAUNCH_DIRECT;
    return VX_OK;
}

extern "C" int32_t vx_init(void) {
    vx::clear_error();
    return vx::ensure_init();
}
