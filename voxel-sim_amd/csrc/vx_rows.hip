// Drop-in device solve over the reference row packing: vx_solve_pcg (f32
// vectors, f64 reductions, f32 alpha/beta, residual check every 16
// iterations — the semantics of src/cuda_solver.rs:301-501 and the kernels of
// src/cuda/pcg.ptx) and vx_solve_pcg_f64 (the same algorithm in FP64).
//
// MI355X design: the caller's array-of-structs rows are uploaded once per
// solve (the boundary hands over host slices) and re-laid out on the device
// into structure-of-arrays planes, so every PCG kernel streams coalesced
// 256-byte wave loads. One iteration is two kernels:
//   rows_spmv   : re-sums the previous update's r.z / r.r partials, checks
//                 convergence, forms beta, then p = z + beta*p_old on load
//                 (own row and neighbour rows), Ap = A p and per-workgroup
//                 f64 partials of p.Ap;
//   rows_update : re-sums the p.Ap partials into alpha, dx += alpha p,
//                 r -= alpha Ap, z = Dinv r, partials of r.z and r.r.
// Scalars are rebuilt from partials inside each kernel (no atomics, no
// fences); workgroup 0 records them in the device state for the next kernel.
// Sixteen iterations are captured once into a hipGraph and replayed; the
// host reads a pinned status word one graph behind, so the device never
// waits for the host.
#include "vx_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace vx {
namespace {

constexpr uint32_t kChunk = 16;  // iterations per graph replay = reference check interval

// Device-resident PCG scalars. Each field is written by workgroup 0 of one
// kernel and read by the following kernels only.
struct RowsState {
    double rho[2];      // r.z after k updates, slot k & 1 (written by rows_spmv)
    double rr_checked;  // r.r at the last convergence check
    double initial;     // ||rhs||
    double tol;         // tolerance in the vector type
    double rel;         // relative residual at the last check
    uint32_t k_spmv;    // updates completed (written by rows_update)
    uint32_t k_update;  // iteration of the current direction (written by rows_spmv)
    uint32_t iters, done, max_iter, check_every;
    uint32_t invalid, pad;
    double red[3];      // pre-summed partials of a big grid (rows_reduce), by slot
};
enum : uint32_t { kRunning = 0, kConverged = 1, kNonFinite = 2, kCapReached = 3 };
enum : int { kPap = 0, kRz = 1, kRr = 2 };  // partial slots

// Grids of more workgroups than this have their partials summed once per
// kernel by rows_reduce instead of by every workgroup of the consumer kernel,
// whose cost grows with the square of the grid. Both sum in the same order
// (thread-strided, then the fixed workgroup sum), so the scalars are the same
// bits either way.
constexpr uint32_t kRowsBig = 1024;

template <int N>
__device__ __forceinline__ void rows_sums(double (&out)[N], const double *partials, const int (&slot)[N], uint32_t nb,
                                          const RowsState *st, double (*lds)[kWaves]) {
    if (nb > kRowsBig) {
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = st->red[slot[k]];
        return;
    }
    sum_partials<N>(out, partials, slot, nb, nb, lds);
}

// One workgroup: slots [k0, k0 + nk) of a big grid's partials into red[].
__global__ __launch_bounds__(kBlock) void rows_reduce(const double *partials, uint32_t nb, int k0, int nk,
                                                      RowsState *st) {
    __shared__ double lds[2][kWaves];
    double s[2] = {0.0, 0.0};
    for (int k = 0; k < nk; ++k)
        for (uint32_t b = threadIdx.x; b < nb; b += kBlock) s[k] += partials[size_t(k0 + k) * nb + b];
    block_sum_all<2>(s, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < nk; ++k) st->red[k0 + k] = s[k];
}

template <class S> __device__ __forceinline__ S fma_s(S a, S b, S c);
template <> __device__ __forceinline__ float fma_s<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fma_s<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Guarded quotient of scalar_ratio / dot_reduce_ratio_single: num/den cast to
// the vector type when den > 1e-30, else 0.
template <class S> __device__ __forceinline__ double guarded_ratio(double num, double den) {
    return den > 1e-30 ? double(S(num / den)) : 0.0;
}

// Per-row dot product with f64 products of the widened elements.
template <class S> __device__ __forceinline__ double row_dot(const S (&a)[3], const S (&b)[3]) {
    double v = double(a[0]) * double(b[0]);
    v = __builtin_fma(double(a[1]), double(b[1]), v);
    v = __builtin_fma(double(a[2]), double(b[2]), v);
    return v;
}

// Inverse of a row-major 3x3 by cofactors, in double.
__device__ __forceinline__ void inverse3(const double (&a)[9], double (&o)[9]) {
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[5] * a[6] - a[3] * a[8];
    double c02 = a[3] * a[7] - a[4] * a[6];
    double inv = 1.0 / (a[0] * c00 + a[1] * c01 + a[2] * c02);
    o[0] = c00 * inv;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * inv;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * inv;
    o[3] = c01 * inv;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * inv;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * inv;
    o[6] = c02 * inv;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * inv;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * inv;
}

// AoS rows -> SoA planes. Plane q of an array with Q values per row lives at
// base + q*n. Neighbour indices that are neither EMPTY nor < n are flagged
// and treated as absent.
template <class S>
__global__ __launch_bounds__(kBlock) void rows_relayout(uint32_t n, const uint32_t *__restrict__ nbr_in,
                                                        const float *__restrict__ blk_in,
                                                        const float *__restrict__ diag_in,
                                                        const float *__restrict__ inv_in,
                                                        const float *__restrict__ rhs_in, uint32_t *__restrict__ nbr,
                                                        S *__restrict__ blk, S *__restrict__ diag, S *__restrict__ inv,
                                                        S *__restrict__ r, S *__restrict__ dx, S *__restrict__ p0,
                                                        RowsState *st, int widen_inverse) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const size_t N = n;
    for (int s = 0; s < 6; ++s) {
        uint32_t j = nbr_in[6 * size_t(i) + s];
        if (j != VX_EMPTY && j >= n) {
            st->invalid = 1;
            j = VX_EMPTY;
        }
        nbr[s * N + i] = j;
    }
    for (int q = 0; q < 54; ++q) blk[q * N + i] = S(blk_in[54 * size_t(i) + q]);
    double d[9];
    for (int q = 0; q < 9; ++q) {
        float v = diag_in[9 * size_t(i) + q];
        diag[q * N + i] = S(v);
        d[q] = double(S(v));
    }
    if (widen_inverse) {
        for (int q = 0; q < 9; ++q) inv[q * N + i] = S(inv_in[9 * size_t(i) + q]);
    } else {
        double o[9];
        inverse3(d, o);
        for (int q = 0; q < 9; ++q) inv[q * N + i] = S(o[q]);
    }
    for (int c = 0; c < 3; ++c) {
        r[c * N + i] = S(rhs_in[3 * size_t(i) + c]);
        dx[c * N + i] = S(0);
        p0[c * N + i] = S(0);
    }
}

template <class S>
__device__ __forceinline__ void apply_inv(const S *__restrict__ inv, size_t N, uint32_t i, const S (&r)[3], S (&z)[3]) {
#pragma unroll
    for (int row = 0; row < 3; ++row) {
        S v = inv[(3 * row) * N + i] * r[0];
        v = fma_s<S>(inv[(3 * row + 1) * N + i], r[1], v);
        v = fma_s<S>(inv[(3 * row + 2) * N + i], r[2], v);
        z[row] = v;
    }
}

// z = Dinv r; partials of r.z and r.r for the first direction.
template <class S>
__global__ __launch_bounds__(kBlock) void rows_init(uint32_t n, uint32_t per_xcd, uint32_t ntiles,
                                                    const S *__restrict__ inv, const S *__restrict__ r,
                                                    S *__restrict__ z, double *partials) {
    __shared__ double lds[2][kWaves];
    const size_t N = n;
    const uint32_t tile = xcd_tile(blockIdx.x, per_xcd);
    const uint32_t i = tile * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (tile < ntiles && i < n) {
        S ri[3] = {r[i], r[N + i], r[2 * N + i]};
        S zi[3];
        apply_inv<S>(inv, N, i, ri, zi);
        z[i] = zi[0];
        z[N + i] = zi[1];
        z[2 * N + i] = zi[2];
        v[0] = row_dot<S>(ri, zi);
        v[1] = row_dot<S>(ri, ri);
    }
    const int slots[2] = {kRz, kRr};
    store_partials<2>(v, partials, slots, gridDim.x, lds);
}

// Convergence check of the previous update (every check_every iterations and
// at the cap, `<=`, src/cuda_solver.rs:458-477), beta = rho_k / rho_{k-1},
// then Ap = A p with p = z + beta*p_old formed on load; partials of p.Ap.
template <class S>
__global__ __launch_bounds__(kBlock) void rows_spmv(uint32_t n, uint32_t per_xcd, uint32_t ntiles,
                                                    const uint32_t *__restrict__ nbr, const S *__restrict__ blk,
                                                    const S *__restrict__ diag, const S *__restrict__ z,
                                                    const S *__restrict__ p_old, S *__restrict__ p_new,
                                                    S *__restrict__ ap, RowsState *st, double *partials) {
    __shared__ double lds[2][kWaves];
    // Scalars first: state words and the previous update's partials load
    // together and are summed across the workgroup.
    const uint32_t done_in = st->done;
    const uint32_t k = st->k_spmv;
    const uint32_t max_iter = st->max_iter, check_every = st->check_every;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const double initial = st->initial, tol = st->tol;
    const uint32_t nb = gridDim.x;
    double sums[2];
    const int in_slots[2] = {kRz, kRr};
    rows_sums<2>(sums, partials, in_slots, nb, st, lds);
    if (done_in) return;
    const size_t N = n;
    const uint32_t tile = xcd_tile(blockIdx.x, per_xcd);
    const uint32_t i = tile * kBlock + threadIdx.x;
    const bool row = tile < ntiles && i < n;
    // Convergence check of the previous update (every check_every iterations
    // and at the cap, `<=`, src/cuda_solver.rs:458-477); beta = rho_k / rho_{k-1}.
    const double rho = sums[0], rr = sums[1];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    uint32_t done = kRunning;
    if (k > 0 && ((check_every && k % check_every == 0) || k == max_iter)) {
        if (!isfinite(rr) || rr < 0.0) {
            done = kNonFinite;
        } else {
            const S rel = S(sqrt(rr) / initial);
            if (lead) st->rel = double(rel);
            if (double(rel) <= tol) done = kConverged;
        }
        if (lead) st->rr_checked = rr;
    }
    if (!done && k >= max_iter) done = kCapReached;
    if (done) {
        if (lead) {
            st->done = done;
            st->iters = k;
        }
        return;
    }
    const S beta = k == 0 ? S(0) : S(guarded_ratio<S>(rho, rho_s[(k - 1) & 1]));
    if (lead) {
        st->rho[k & 1] = rho;
        st->k_update = k;
    }
    // p = z + beta p_old on load, y = D p + sum_s B_s p_s in the reference
    // kernel's fused multiply-add order. All loads are issued before any
    // arithmetic (absent slots read the row itself and are skipped).
    double v[1] = {0.0};
    if (row) {
        uint32_t jj[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) jj[s] = nbr[s * N + i];
        S zz[7][3], pp[7][3], dg[9];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            zz[0][c] = z[c * N + i];
            pp[0][c] = p_old[c * N + i];
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const uint32_t j = jj[s] == VX_EMPTY ? i : jj[s];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                zz[1 + s][c] = z[c * N + j];
                pp[1 + s][c] = p_old[c * N + j];
            }
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) dg[q] = diag[q * N + i];
        S bb[54];
#pragma unroll
        for (int q = 0; q < 54; ++q) bb[q] = blk[q * N + i];
        S p[3], y[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = fma_s<S>(beta, pp[0][c], zz[0][c]);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            S a = dg[3 * r] * p[0];
            a = fma_s<S>(dg[3 * r + 1], p[1], a);
            a = fma_s<S>(dg[3 * r + 2], p[2], a);
            y[r] = a;
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            if (jj[s] == VX_EMPTY) continue;
            S pj[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) pj[c] = fma_s<S>(beta, pp[1 + s][c], zz[1 + s][c]);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const S *b = bb + 9 * s + 3 * r;
                y[r] = fma_s<S>(b[0], pj[0], y[r]);
                y[r] = fma_s<S>(b[1], pj[1], y[r]);
                y[r] = fma_s<S>(b[2], pj[2], y[r]);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            p_new[c * N + i] = p[c];
            ap[c * N + i] = y[c];
        }
        v[0] = row_dot<S>(p, y);
    }
    const int out_slots[1] = {kPap};
    store_partials<1>(v, partials, out_slots, nb, lds);
}

// alpha = rho / p.Ap (guarded); dx += alpha p; r -= alpha Ap; z = Dinv r;
// partials of r.z and r.r.
template <class S>
__global__ __launch_bounds__(kBlock) void rows_update(uint32_t n, uint32_t per_xcd, uint32_t ntiles,
                                                      const S *__restrict__ inv, S *__restrict__ dx,
                                                      S *__restrict__ r, S *__restrict__ z,
                                                      const S *__restrict__ p, const S *__restrict__ ap,
                                                      RowsState *st, double *partials) {
    __shared__ double lds[2][kWaves];
    const uint32_t done_in = st->done;
    const uint32_t k = st->k_update;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const uint32_t nb = gridDim.x;
    double pap[1];
    const int in_slots[1] = {kPap};
    rows_sums<1>(pap, partials, in_slots, nb, st, lds);
    if (done_in) return;
    const size_t N = n;
    const uint32_t tile = xcd_tile(blockIdx.x, per_xcd);
    const uint32_t i = tile * kBlock + threadIdx.x;
    const bool row = tile < ntiles && i < n;
    S pv[3], av[3], rv[3], dv[3];
    if (row) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pv[c] = p[c * N + i];
            av[c] = ap[c * N + i];
            rv[c] = r[c * N + i];
            dv[c] = dx[c * N + i];
        }
    }
    const S alpha = S(guarded_ratio<S>(rho_s[k & 1], pap[0]));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->k_spmv = k + 1;
        st->iters = k + 1;
    }
    double v[2] = {0.0, 0.0};
    if (row) {
        S ri[3], zi[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dx[c * N + i] = fma_s<S>(alpha, pv[c], dv[c]);
            ri[c] = fma_s<S>(-alpha, av[c], rv[c]);
            r[c * N + i] = ri[c];
        }
        apply_inv<S>(inv, N, i, ri, zi);
#pragma unroll
        for (int c = 0; c < 3; ++c) z[c * N + i] = zi[c];
        v[0] = row_dot<S>(ri, zi);
        v[1] = row_dot<S>(ri, ri);
    }
    const int out_slots[2] = {kRz, kRr};
    store_partials<2>(v, partials, out_slots, nb, lds);
}

// SoA correction -> caller's AoS f32 layout.
template <class S>
__global__ __launch_bounds__(kBlock) void rows_pack_out(uint32_t n, const S *__restrict__ dx, float *__restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const size_t N = n;
    for (int c = 0; c < 3; ++c) out[3 * size_t(i) + c] = float(dx[c * N + i]);
}

// One captured 16-iteration graph per (precision, size, buffers) key.
struct RowsGraph {
    hipGraphExec_t exec = nullptr;
    const void *key[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t n = 0;
};
RowsGraph g_graph[2];

// Scratch area of vx_solve_pcg_noalloc: every device buffer of one f32 row
// solve (staging copies of the caller's rows, the SoA planes, the vectors,
// partials, state and output), carved out of one statically allocated array
// per device. Host code never passes `scratch` itself to a kernel or a copy
// (in host code it names a host-side stand-in): it uses the device address
// obtained once per device with hipGetSymbolAddress.
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB, held by each device until the process ends
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
void *g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of `scratch` on the current device.
int32_t scratch_address(void **out) {
    int dev = 0;
    VX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(VX_ERR_DEVICE, "device %d outside the scratch table", dev);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(VX_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return VX_OK;
}

}  // namespace

// Called by the library's initialisation with device 0 current: looks the
// scratch area's device address up once, so the allocation-free entry point
// does no set-up of its own (it retries if this failed).
void rows_scratch_setup() {
    void *p = nullptr;
    (void)scratch_address(&p);
}

namespace {

// Bytes of the solve's buffers for n rows and a grid of `grid` workgroups,
// each rounded up to 256 (the scratch carve-out's alignment).
template <class S>
size_t rows_buffer_bytes(size_t N, uint32_t grid, size_t state_bytes) {
    const size_t sizes[] = {6 * N * 4, 54 * N * 4, 9 * N * 4, 9 * N * 4, 3 * N * 4,  // staging
                            6 * N * 4, 54 * N * sizeof(S), 9 * N * sizeof(S), 9 * N * sizeof(S),
                            3 * N * sizeof(S), 3 * N * sizeof(S), 3 * N * sizeof(S), 3 * N * sizeof(S),
                            3 * N * sizeof(S), 3 * N * sizeof(S), 4 * size_t(grid) * 8, state_bytes, 3 * N * 4};
    size_t total = 0;
    for (size_t b : sizes) total += (std::max<size_t>(b, 256) + 255) & ~size_t(255);
    return total;
}

// noalloc: the buffers come from the scratch area, the chunks are launched
// directly (no graph instantiation) and the host waits for each chunk with a
// stream synchronisation (no events), so the call allocates nothing.
template <class S>
int32_t solve_rows(const float *rhs, const uint32_t *neighbors, const float *diagonal, const float *blocks,
                   const float *inverse_diagonal, uint32_t n, uint32_t max_iterations, float tolerance,
                   float *out_correction, uint32_t *out_iterations, float *out_relative_residual,
                   bool noalloc = false) {
    constexpr bool kF64 = sizeof(S) == 8;
    clear_error();
    if (n > 0 && (!rhs || !neighbors || !diagonal || !blocks || !inverse_diagonal || !out_correction))
        return fail(VX_ERR_INVALID_ARGUMENT, "null array with n=%u", n);
    if (n == 0) {
        if (out_iterations) *out_iterations = 0;
        if (out_relative_residual) *out_relative_residual = 0.0f;
        return VX_OK;
    }
    // ||rhs|| in f64 on the host (src/cuda_solver.rs:316-324).
    double initial_sq = 0.0;
    for (size_t k = 0; k < 3 * size_t(n); ++k) initial_sq += double(rhs[k]) * double(rhs[k]);
    const double initial = std::sqrt(initial_sq);
    if (initial <= 1e-12) {
        std::memset(out_correction, 0, sizeof(float) * 3 * size_t(n));
        if (out_iterations) *out_iterations = 0;
        if (out_relative_residual) *out_relative_residual = 0.0f;
        return VX_OK;
    }
    const size_t N = n;
    uint32_t per_xcd = 0;
    const uint32_t ntiles = (n + kBlock - 1) / kBlock;
    const uint32_t grid = xcd_grid(ntiles, &per_xcd);
    if (noalloc) {
        const size_t need = rows_buffer_bytes<S>(N, grid, sizeof(RowsState));
        if (need > kScratchBytes)
            return fail(VX_ERR_TOO_LARGE, "%u rows need %zu bytes of scratch; the allocation-free solve has %zu", n, need,
                        kScratchBytes);
    }
    int32_t st_code = ensure_init();
    if (st_code != VX_OK) return st_code;
    Context &ctx = context();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex);
    hipStream_t stream = ctx.stream;
    VX_HIP(hipSetDevice(0));
    char *scratch_base = nullptr;
    size_t scratch_used = 0;
    if (noalloc)
        if (int32_t c = scratch_address(reinterpret_cast<void **>(&scratch_base))) return c;

    void *s_nbr, *s_blk, *s_diag, *s_inv, *s_rhs;
    void *d_nbr, *d_blk, *d_diag, *d_inv, *d_dx, *d_r, *d_z, *d_p0, *d_p1, *d_ap, *d_part, *d_state, *d_out;
    const char *pre = kF64 ? "rows64_" : "rows32_";
    auto B = [&](const char *name, size_t bytes, void **out) -> int32_t {
        if (noalloc) {  // carve out of the scratch area (size checked above)
            *out = scratch_base + scratch_used;
            scratch_used += (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
            return VX_OK;
        }
        std::string key = std::string(pre) + name;
        return buffer(key.c_str(), bytes, out);
    };
#define VX_BUF(name, bytes, ptr)                      \
    do {                                              \
        int32_t c_ = B(name, bytes, ptr);             \
        if (c_ != VX_OK) return c_;                   \
    } while (0)
    VX_BUF("stage_nbr", 6 * N * 4, &s_nbr);
    VX_BUF("stage_blk", 54 * N * 4, &s_blk);
    VX_BUF("stage_diag", 9 * N * 4, &s_diag);
    VX_BUF("stage_inv", 9 * N * 4, &s_inv);
    VX_BUF("stage_rhs", 3 * N * 4, &s_rhs);
    VX_BUF("nbr", 6 * N * 4, &d_nbr);
    VX_BUF("blk", 54 * N * sizeof(S), &d_blk);
    VX_BUF("diag", 9 * N * sizeof(S), &d_diag);
    VX_BUF("inv", 9 * N * sizeof(S), &d_inv);
    VX_BUF("dx", 3 * N * sizeof(S), &d_dx);
    VX_BUF("r", 3 * N * sizeof(S), &d_r);
    VX_BUF("z", 3 * N * sizeof(S), &d_z);
    VX_BUF("p0", 3 * N * sizeof(S), &d_p0);
    VX_BUF("p1", 3 * N * sizeof(S), &d_p1);
    VX_BUF("ap", 3 * N * sizeof(S), &d_ap);
    VX_BUF("partials", 4 * size_t(grid) * 8, &d_part);
    VX_BUF("state", sizeof(RowsState), &d_state);
    VX_BUF("out", 3 * N * 4, &d_out);
#undef VX_BUF

    VX_HIP(hipMemcpyAsync(s_nbr, neighbors, 6 * N * 4, hipMemcpyHostToDevice, stream));
    VX_HIP(hipMemcpyAsync(s_blk, blocks, 54 * N * 4, hipMemcpyHostToDevice, stream));
    VX_HIP(hipMemcpyAsync(s_diag, diagonal, 9 * N * 4, hipMemcpyHostToDevice, stream));
    VX_HIP(hipMemcpyAsync(s_inv, inverse_diagonal, 9 * N * 4, hipMemcpyHostToDevice, stream));
    VX_HIP(hipMemcpyAsync(s_rhs, rhs, 3 * N * 4, hipMemcpyHostToDevice, stream));

    RowsState hs;
    std::memset(&hs, 0, sizeof hs);
    hs.initial = initial;
    hs.tol = kF64 ? double(tolerance) : double(float(tolerance));
    hs.rel = INFINITY;
    hs.rr_checked = INFINITY;
    hs.k_spmv = 0;
    hs.k_update = 0;
    hs.max_iter = max_iterations;
    hs.check_every = kChunk;
    hs.done = max_iterations == 0 ? kCapReached : kRunning;
    RowsState *pinned = static_cast<RowsState *>(ctx.pinned);
    std::memcpy(&pinned[2], &hs, sizeof hs);
    VX_HIP(hipMemcpyAsync(d_state, &pinned[2], sizeof hs, hipMemcpyHostToDevice, stream));

    RowsState *st = static_cast<RowsState *>(d_state);
    S *dx = static_cast<S *>(d_dx), *r = static_cast<S *>(d_r), *z = static_cast<S *>(d_z);
    S *p0 = static_cast<S *>(d_p0), *p1 = static_cast<S *>(d_p1), *ap = static_cast<S *>(d_ap);
    S *blk = static_cast<S *>(d_blk), *diag = static_cast<S *>(d_diag), *inv = static_cast<S *>(d_inv);
    uint32_t *nbr = static_cast<uint32_t *>(d_nbr);
    double *part = static_cast<double *>(d_part);

    const uint32_t rgrid = ntiles;
    hipLaunchKernelGGL(rows_relayout<S>, dim3(rgrid), dim3(kBlock), 0, stream, n,
                       static_cast<const uint32_t *>(s_nbr), static_cast<const float *>(s_blk),
                       static_cast<const float *>(s_diag), static_cast<const float *>(s_inv),
                       static_cast<const float *>(s_rhs), nbr, blk, diag, inv, r, dx, p0, st, kF64 ? 0 : 1);
    VX_HIP(hipGetLastError());
    if (max_iterations > 0) {
        hipLaunchKernelGGL(rows_init<S>, dim3(grid), dim3(kBlock), 0, stream, n, per_xcd, ntiles, inv, r, z, part);
        const bool big = grid > kRowsBig;
        if (big) hipLaunchKernelGGL(rows_reduce, dim3(1), dim3(kBlock), 0, stream, part, grid, int(kRz), 2, st);
        VX_HIP(hipGetLastError());

        // One chunk of kChunk iterations: captured once into a graph and
        // replayed, or launched kernel by kernel in direct launch mode.
        auto enqueue_chunk = [&]() {
            for (uint32_t t = 0; t < kChunk; ++t) {
                S *pold = (t % 2 == 0) ? p0 : p1;
                S *pnew = (t % 2 == 0) ? p1 : p0;
                hipLaunchKernelGGL(rows_spmv<S>, dim3(grid), dim3(kBlock), 0, stream, n, per_xcd, ntiles, nbr, blk,
                                   diag, z, pold, pnew, ap, st, part);
                if (big) hipLaunchKernelGGL(rows_reduce, dim3(1), dim3(kBlock), 0, stream, part, grid, int(kPap), 1, st);
                hipLaunchKernelGGL(rows_update<S>, dim3(grid), dim3(kBlock), 0, stream, n, per_xcd, ntiles, inv, dx,
                                   r, z, pnew, ap, st, part);
                if (big) hipLaunchKernelGGL(rows_reduce, dim3(1), dim3(kBlock), 0, stream, part, grid, int(kRz), 2, st);
            }
        };
        const bool direct = ctx.direct_launch || noalloc;
        RowsGraph &g = g_graph[kF64 ? 1 : 0];
        const void *key[4] = {d_blk, d_p0, d_part, d_state};
        bool reuse = direct || (g.exec && g.n == n && std::memcmp(key, g.key, sizeof key) == 0);
        if (!reuse) {
            if (g.exec) {
                VX_HIP(hipGraphExecDestroy(g.exec));
                g.exec = nullptr;
            }
            hipGraph_t graph;
            VX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            enqueue_chunk();
            VX_HIP(hipStreamEndCapture(stream, &graph));
            hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(VX_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            g.n = n;
            std::memcpy(g.key, key, sizeof key);
        }
        uint32_t launched = 0, chunk = 0;
        int32_t rc = VX_OK;
        if (noalloc) {
            // One chunk at a time, the status read after each (no events):
            // the chunks a lookahead would issue after convergence change
            // nothing, so the result is the same.
            while (launched <= max_iterations) {
                enqueue_chunk();
                hipError_t e = hipGetLastError();
                if (e == hipSuccess) e = hipMemcpyAsync(&pinned[0], d_state, sizeof hs, hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipStreamSynchronize(stream);
                if (e != hipSuccess) {
                    rc = fail(VX_ERR_DEVICE, "PCG chunk failed: %s", hipGetErrorString(e));
                    break;
                }
                launched += kChunk;
                if (pinned[0].done) break;
            }
        } else {
            hipEvent_t ev[2];
            VX_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
            VX_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
            // The check after the last update runs in the following rows_spmv, so
            // replay until one spmv past the cap has been issued.
            while (launched <= max_iterations) {
                hipError_t e = hipSuccess;
                if (direct) {
                    enqueue_chunk();
                    e = hipGetLastError();
                } else {
                    e = hipGraphLaunch(g.exec, stream);
                }
                if (e == hipSuccess) e = hipMemcpyAsync(&pinned[chunk % 2], d_state, sizeof hs, hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipEventRecord(ev[chunk % 2], stream);
                if (e != hipSuccess) {
                    rc = fail(VX_ERR_DEVICE, "PCG graph launch failed: %s", hipGetErrorString(e));
                    break;
                }
                launched += kChunk;
                if (chunk >= 1) {
                    e = hipEventSynchronize(ev[(chunk - 1) % 2]);
                    if (e != hipSuccess) {
                        rc = fail(VX_ERR_DEVICE, "PCG status read failed: %s", hipGetErrorString(e));
                        break;
                    }
                    if (pinned[(chunk - 1) % 2].done) break;
                }
                chunk += 1;
            }
            (void)hipStreamSynchronize(stream);
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
        }
        if (rc != VX_OK) return rc;
    }
    hipLaunchKernelGGL(rows_pack_out<S>, dim3(rgrid), dim3(kBlock), 0, stream, n, dx, static_cast<float *>(d_out));
    VX_HIP(hipGetLastError());
    VX_HIP(hipMemcpyAsync(&pinned[2], d_state, sizeof hs, hipMemcpyDeviceToHost, stream));
    VX_HIP(hipStreamSynchronize(stream));
    const RowsState fin = pinned[2];
    if (fin.invalid) return fail(VX_ERR_INVALID_ARGUMENT, "neighbour index out of range (n=%u)", n);
    if (out_iterations) *out_iterations = fin.iters;
    if (out_relative_residual) *out_relative_residual = float(fin.rel);
    const char *dev = "HIP PCG";
    if (fin.done == kNonFinite || !std::isfinite(fin.rr_checked) || fin.rr_checked < 0.0)
        return fail(VX_ERR_NONFINITE_RESIDUAL, "%s produced a non-finite residual norm", dev);
    if (!std::isfinite(fin.rel) || fin.rel > fin.tol)
        return fail(VX_ERR_NOT_CONVERGED,
                    "%s did not converge: iterations=%u, relative residual=%.3e, tolerance=%.3e", dev, fin.iters,
                    fin.rel, double(tolerance));
    VX_HIP(hipMemcpy(out_correction, d_out, 3 * N * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < 3 * N; ++k)
        if (!std::isfinite(out_correction[k])) return fail(VX_ERR_NONFINITE_CORRECTION, "%s produced a non-finite correction", dev);
    return VX_OK;
}

}  // namespace
}  // namespace vx

extern "C" int32_t vx_solve_pcg(const float *rhs, const uint32_t *neighbors, const float *diagonal, const float *blocks,
                                const float *inverse_diagonal, uint32_t n, uint32_t max_iterations, float tolerance,
                                float *out_correction, uint32_t *out_iterations, float *out_relative_residual) {
    return vx::solve_rows<float>(rhs, neighbors, diagonal, blocks, inverse_diagonal, n, max_iterations, tolerance,
                                 out_correction, out_iterations, out_relative_residual);
}

extern "C" int32_t vx_solve_pcg_noalloc(const float *rhs, const uint32_t *neighbors, const float *diagonal,
                                        const float *blocks, const float *inverse_diagonal, uint32_t n,
                                        uint32_t max_iterations, float tolerance, float *out_correction,
                                        uint32_t *out_iterations, float *out_relative_residual) {
    return vx::solve_rows<float>(rhs, neighbors, diagonal, blocks, inverse_diagonal, n, max_iterations, tolerance,
                                 out_correction, out_iterations, out_relative_residual, true);
}

extern "C" int32_t vx_solve_pcg_f64(const float *rhs, const uint32_t *neighbors, const float *diagonal,
                                    const float *blocks, const float *inverse_diagonal, uint32_t n,
                                    uint32_t max_iterations, float tolerance, float *out_correction,
                                    uint32_t *out_iterations, float *out_relative_residual) {
    return vx::solve_rows<double>(rhs, neighbors, diagonal, blocks, inverse_diagonal, n, max_iterations, tolerance,
                                  out_correction, out_iterations, out_relative_residual);
}
