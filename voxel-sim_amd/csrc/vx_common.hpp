// Shared runtime pieces of the gfx950 PCG library: the process-wide device
// context, thread-local error messages, grow-only device buffers and the
// wave64 / workgroup reduction helpers used by every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "../../include/vx_solver.h"

namespace vx {

constexpr int kBlock = 512;     // 8 wave64s per workgroup
constexpr int kWaves = kBlock / 64;
constexpr int kXcds = 8;        // MI355X: 8 XCDs, workgroups dispatched round-robin

// ---- errors -----------------------------------------------------------------
int32_t fail(int32_t code, const char *fmt, ...);
void clear_error();

#define VX_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            return ::vx::fail(e_ == hipErrorOutOfMemory ? VX_ERR_TOO_LARGE : VX_ERR_DEVICE, \
                              "%s failed: %s", #call, hipGetErrorString(e_));               \
        }                                                                                   \
    } while (0)

// ---- device context ----------------------------------------------------------
// One context per process on device 0, serialised by a mutex, as the
// reference's OnceLock<Mutex<CudaSolver>> (src/cuda_solver.rs:66, 326-336).
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct Context {
    bool initialised = false;
    int32_t init_status = VX_OK;
    std::string init_message;
    hipStream_t stream = nullptr;
    std::recursive_mutex mutex;
    std::map<std::string, DeviceBuffer> buffers;  // named, grow-only (src/cuda_solver.rs:162-193)
    void *pinned = nullptr;                       // small pinned host area for status words
    size_t pinned_bytes = 0;
    // Launch the PCG chunks kernel by kernel instead of replaying their
    // captured hipGraphs (vx_set_launch_mode; the same kernels, arguments and
    // order).
    bool direct_launch = false;
    // How vx_contact_project runs its passes (vx_contact_set_method).
    int32_t contact_method = VX_CONTACT_AUTO;
    uint32_t contact_rounds = 1536;  // relaxation rounds per call before the passes left go to windows
};

Context &context();
// Initialise once; a failure is cached for the process lifetime.
int32_t ensure_init();
// Named grow-only device buffer of at least `bytes`.
int32_t buffer(const char *name, size_t bytes, void **out);
// Device address of the allocation-free row solve's scratch area, looked up
// at initialisation (vx_rows.hip).
void rows_scratch_setup();

// XCD-aware tile index: consecutive logical tiles land on the same XCD, so
// the z-neighbour rows a tile reads were fetched into that XCD's L2 by the
// tiles beside it. Hardware assigns workgroup b to XCD b % 8.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t per_xcd) {
    return (b % kXcds) * per_xcd + b / kXcds;
}
inline uint32_t xcd_grid(uint32_t tiles, uint32_t *per_xcd) {
    *per_xcd = (tiles + kXcds - 1) / kXcds;
    return *per_xcd * kXcds;
}

// ---- row loads ----------------------------------------------------------------
// Element access through a 32-bit byte offset from a plane's base, which must
// be the same for every lane (a plane of an SoA field). The base is pinned to
// scalar registers (readfirstlane), so the compiler emits the SGPR-base form
// of global_load/store (one VGPR offset shared by every plane of a row)
// instead of a 64-bit address add per access. Offsets stay below 2^32:
// arenas hold fewer than 2^29 rows.
#define VX_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ uint64_t uniform_addr(const void *p) {
    const uint64_t x = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
    return (uint64_t(hi) << 32) | lo;
}
template <class T>
__device__ __forceinline__ T ld_row(const T *plane, uint32_t byte_off) {
    const VX_GLOBAL char *base = reinterpret_cast<const VX_GLOBAL char *>(uniform_addr(plane));
    return *reinterpret_cast<const VX_GLOBAL T *>(base + byte_off);
}
template <class T>
__device__ __forceinline__ void st_row(T *plane, uint32_t byte_off, T value) {
    VX_GLOBAL char *base = reinterpret_cast<VX_GLOBAL char *>(uniform_addr(plane));
    *reinterpret_cast<VX_GLOBAL T *>(base + byte_off) = value;
}
// Streaming (nontemporal) forms of the row access, for planes a launch reads
// once or writes for a later kernel, when the arena's planes are too large
// to stay in the caches between kernels anyway: their lines then do not
// displace the neighbour planes that the same launch re-reads.
template <class T>
__device__ __forceinline__ T ld_row_nt(const T *plane, uint32_t byte_off) {
    const VX_GLOBAL char *base = reinterpret_cast<const VX_GLOBAL char *>(uniform_addr(plane));
    return __builtin_nontemporal_load(reinterpret_cast<const VX_GLOBAL T *>(base + byte_off));
}
template <class T>
__device__ __forceinline__ void st_row_nt(T *plane, uint32_t byte_off, T value) {
    VX_GLOBAL char *base = reinterpret_cast<VX_GLOBAL char *>(uniform_addr(plane));
    __builtin_nontemporal_store(value, reinterpret_cast<VX_GLOBAL T *>(base + byte_off));
}

// ---- reductions -------------------------------------------------------------
// Cross-workgroup sums never use atomics or fences: each kernel writes one
// partial per workgroup with a plain store, and the kernels that need the
// total re-sum the partials in every workgroup (the kernel boundary makes
// them visible). The order of the sum is fixed, so every workgroup gets the
// same bits and results are reproducible run to run.
// One DPP step on a double: the two 32-bit halves are moved with the same
// DPP control (lanes without a source, or masked off, read 0) and added.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double dpp_shifted(double v) {
    const uint64_t x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_update_dpp(0u, uint32_t(x), CTRL, ROW_MASK, BANK_MASK, true);
    const uint32_t hi = __builtin_amdgcn_update_dpp(0u, uint32_t(x >> 32), CTRL, ROW_MASK, BANK_MASK, true);
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// Sum over the 64 lanes of a wave, returned to every lane. DPP row shifts
// and row broadcasts stay in the VALU (a shuffle butterfly would make six
// dependent LDS round trips); the order is fixed, so the bits are
// reproducible. Steps: lane i of each 16-lane row gathers lanes i-3..i of the
// original values (row_shr 1, 2, 3), then i-7..i (row_shr 4 on banks 1-3) and
// the row prefix (row_shr 8 on banks 2-3); row_bcast 15 and 31 carry the row
// totals upward, and lane 63 ends with the wave total.
__device__ __forceinline__ double wave_sum(double v) {
    double s = v + dpp_shifted<0x111, 0xf, 0xf>(v);  // row_shr:1
    s += dpp_shifted<0x112, 0xf, 0xf>(v);            // row_shr:2
    s += dpp_shifted<0x113, 0xf, 0xf>(v);            // row_shr:3
    s += dpp_shifted<0x114, 0xf, 0xe>(s);            // row_shr:4, banks 1-3
    s += dpp_shifted<0x118, 0xf, 0xc>(s);            // row_shr:8, banks 2-3
    s += dpp_shifted<0x142, 0xa, 0xf>(s);            // row_bcast:15, rows 1, 3
    s += dpp_shifted<0x143, 0xc, 0xf>(s);            // row_bcast:31, rows 2, 3
    const uint64_t x = __builtin_bit_cast(uint64_t, s);
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(x), 63);
    const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(x >> 32), 63);
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Workgroup barrier that orders LDS only. Cross-thread data in these
// reductions moves through LDS alone, so the barrier waits for LDS traffic
// (lgkmcnt) but not for global loads or stores still in flight (a full
// __syncthreads() would wait for vmcnt(0) and serialise the row loads issued
// before it).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroup sum of N values, broadcast to every thread.
template <int N>
__device__ __forceinline__ void block_sum_all(double (&v)[N], double (*lds)[kWaves]) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    lds_barrier();  // a previous use of lds is finished
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) lds[k][wave] = s;
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = lds[k][0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += lds[k][w];
        v[k] = s;
    }
}

// Store this workgroup's N partials: partials[slot[k] * stride + blockIdx.x].
template <int N>
__device__ __forceinline__ void store_partials(double (&v)[N], double *partials, const int (&slot)[N], uint32_t stride,
                                               double (*lds)[kWaves]) {
    block_sum_all<N>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) partials[size_t(slot[k]) * stride + blockIdx.x] = v[k];
    }
}

// Sum N partial arrays of `count` entries (written by the previous kernel),
// broadcast to every thread.
template <int N>
__device__ __forceinline__ void sum_partials(double (&out)[N], const double *partials, const int (&slot)[N],
                                             uint32_t stride, uint32_t count, double (*lds)[kWaves]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
        for (uint32_t b = threadIdx.x; b < count; b += kBlock) s += partials[size_t(slot[k]) * stride + b];
        out[k] = s;
    }
    block_sum_all<N>(out, lds);
}

}  // namespace vx
