// Device contact projection: the drop-in for collision::project
// (src/collision.rs:33-123).
//
// The reference projects voxel pairs apart with a Gauss-Seidel sweep over
// the candidate pairs sorted by (i, j), up to six passes, in which every
// `separate` moves positions that later pairs read. Its result depends on
// that order, and this file reproduces it bit for bit:
//
//  * Broad phase. Each voxel's swept box covers a range of unit bins
//    (bin_bounds, src/collision.rs:432-439); two voxels are candidates when
//    their ranges share a bin and they are not both kinematic
//    (src/collision.rs:376-430). Every covered bin becomes an entry (bin code,
//    voxel), the entries are radix sorted by bin, and the voxels of one bin
//    form pairs. A pair is emitted only from the lowest-corner bin of the
//    two ranges' intersection, so each pair appears once without a hash set.
//    Pairs joined by intact bonds inside their lattice cell are dropped
//    (local_continuum_pair, :520-550) and the rest radix sorted by (i, j).
//
//  * Passes by relaxation (the default). A pass's result is the solution of
//    its pair recurrence: pair k evaluates the pass body (swept, overlap and
//    touching tests, :552-621, :441-460) on the positions its voxels have
//    after the moves of the earlier pairs that touch them. Every pair is
//    evaluated on the positions the current outcomes of its earlier pairs
//    give, the pairs whose inputs changed are evaluated again, and so on until
//    no outcome changes: that fixed point is the sequential sweep's result,
//    reached in about as many rounds as the longest chain of moves feeding
//    each other. Only then are the moves applied, per voxel in pair order.
//
//  * Passes by windows (VX_CONTACT_WINDOWS, and any pass the relaxation does
//    not finish within its budget). One workgroup of 1024 threads takes a
//    window of the smallest unprocessed pairs, evaluates every pair on the
//    current positions, and commits the pairs whose sequential predecessors
//    cannot change what they read: a pair is held back when an earlier pair
//    of the window moves one of its voxels, or when an earlier held-back pair
//    shares a voxel that can move (not kinematic). Committed pairs touch
//    disjoint moving voxels, so they commute with the held-back ones and the
//    result equals the sequential sweep. The window's smallest pair always
//    commits.
//
//  * Later passes. A pair whose voxels have not moved since its last
//    evaluation, which did not separate, would repeat that outcome: both
//    methods start a later pass from the other pairs only (the windows check
//    every skipped pair afterwards and redo the pass from a snapshot with
//    more pairs when one was due).
//
//  * Signed zeros. `separate` also adds exact zeros to the components off the
//    normal (and to immovable endpoints). Adding -0 changes nothing and adding
//    +0 only turns a -0 into +0, so those additions commute with everything:
//    the windows run them after the real moves of a window, as a store of +0
//    over -0; the relaxation adds every component as separate() does.
//
// All arithmetic is f32 in the reference's order without contraction (Rust
// never fuses a multiply and an add), with IEEE division.
#include "vx_common.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <cstring>
#include <vector>

#pragma clang fp contract(off)

namespace vx {
namespace {

constexpr float kCell = 0.01f;  // src/scene.rs:10
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kNoMark = 0xFFFFFFFFu;
constexpr float kBinLimit = 1048576.0f;  // |bin| < 2^20: 21 bits per axis in a bin code
constexpr int kBinBias = 1 << 20;
constexpr int kGs = 1024;  // projection workgroup
constexpr int kGsWaves = kGs / 64;
constexpr int kPasses = 6;  // src/collision.rs:55

// flags byte per voxel: bit 0 kinematic, bit 1 + a: axis a held by a target.
__device__ __forceinline__ bool kinematic(uint8_t f) { return f & 1; }

// ---- broad phase -------------------------------------------------------------

// adjacency (src/collision.rs:506-518): slot 2a+1 is the +a neighbour, 2a the -a one.
__global__ void k_adjacency(uint32_t n, const uint32_t *__restrict__ bond, uint32_t *__restrict__ adj,
                            uint32_t *__restrict__ err) {
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= uint64_t(3) * n) return;
    const uint32_t axis = uint32_t(g / n), i = uint32_t(g % n);
    const uint32_t j = bond[g];
    if (j == kEmpty) return;
    if (j >= n) {
        atomicOr(err, 1u);
        return;
    }
    adj[size_t(i) * 6 + 2 * axis + 1] = j;
    adj[size_t(j) * 6 + 2 * axis] = i;
}

// The bin range of each voxel's swept box (src/collision.rs:383-386, 432-439).
// Also the bounding box of every swept box's bins, per workgroup:
// box[6 b + 0..2] the minima (as x + 2^20, so unsigned comparisons order
// them), box[6 b + 3..5] the maxima.
__global__ void k_bins(uint32_t n, const float *__restrict__ start, const float *__restrict__ pos,
                       int32_t *__restrict__ lohi, uint64_t *__restrict__ count, uint32_t *__restrict__ err,
                       uint32_t *__restrict__ box) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bmin[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, bmax[3] = {0, 0, 0};
    if (v < n) {
        const float m = 0.5f + 2e-4f;
        uint64_t c = 1;
        bool ok = true;
        int32_t lo_[3], hi_[3];
        for (int a = 0; a < 3; ++a) {
            const float s = start[size_t(v) * 3 + a], p = pos[size_t(v) * 3 + a];
            const float lo = floorf(fminf(s, p) / kCell - m);
            const float hi = floorf(fmaxf(s, p) / kCell + m);
            // NaN fails both comparisons
            if (!(lo >= -kBinLimit && hi < kBinLimit)) {
                ok = false;
                break;
            }
            lo_[a] = int32_t(lo);
            hi_[a] = int32_t(hi);
            lohi[size_t(v) * 6 + a] = lo_[a];
            lohi[size_t(v) * 6 + 3 + a] = hi_[a];
            c *= uint64_t(int64_t(hi) - int64_t(lo) + 1);
        }
        if (!ok) {
            atomicOr(err, 2u);
            c = 0;
        } else if (c > (uint64_t(1) << 31)) {
            // more entries than the whole call may hold; zeroing it keeps the
            // scan of at most 2^32 counts of at most 2^31 each from wrapping
            atomicOr(err, 4u);
            c = 0;
        }
        if (c) {
            for (int a = 0; a < 3; ++a) {
                bmin[a] = uint32_t(lo_[a] + kBinBias);
                bmax[a] = uint32_t(hi_[a] + kBinBias);
            }
        }
        count[v] = c;
    }
    // workgroup minima and maxima, one partial row per workgroup (k_box_reduce
    // folds them: many atomics on six words would serialise)
    __shared__ uint32_t red[6][4];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            bmin[a] = min(bmin[a], uint32_t(__shfl_xor(int(bmin[a]), off, 64)));
            bmax[a] = max(bmax[a], uint32_t(__shfl_xor(int(bmax[a]), off, 64)));
        }
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int a = 0; a < 3; ++a) {
            red[a][wave] = bmin[a];
            red[3 + a][wave] = bmax[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t q = threadIdx.x;
        uint32_t r = red[q][0];
        for (uint32_t w = 1; w < blockDim.x / 64; ++w) r = q < 3 ? min(r, red[q][w]) : max(r, red[q][w]);
        box[size_t(blockIdx.x) * 6 + q] = r;
    }
}
// Fold the per-workgroup rows of k_bins (minima: 0xFFFFFFFF for a workgroup
// without a box) into out[0..5]; one workgroup.
__global__ void k_box_reduce(uint32_t rows, const uint32_t *__restrict__ part, uint32_t *out) {
    uint32_t r[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < rows; b += blockDim.x)
        for (int q = 0; q < 6; ++q) r[q] = q < 3 ? min(r[q], part[size_t(b) * 6 + q]) : max(r[q], part[size_t(b) * 6 + q]);
    __shared__ uint32_t red[6][16];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = uint32_t(__shfl_xor(int(r[q]), off, 64));
            r[q] = q < 3 ? min(r[q], o) : max(r[q], o);
        }
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = r[q];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t q = threadIdx.x;
        uint32_t x = red[q][0];
        for (uint32_t w = 1; w < blockDim.x / 64; ++w) x = q < 3 ? min(x, red[q][w]) : max(x, red[q][w]);
        out[q] = x;
    }
}
// Sum of the per-workgroup candidate counts of k_bin_pairs; one workgroup.
__global__ void k_sum_reduce(uint32_t rows, const unsigned long long *__restrict__ part, unsigned long long *out) {
    unsigned long long x = 0;
    for (uint32_t b = threadIdx.x; b < rows; b += blockDim.x) x += part[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    __shared__ unsigned long long red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += red[w];
        *out = t;
    }
}

// Bin codes count the bins of the call's bounding box of swept boxes:
// ((z - z0) Y + (y - y0)) X + (x - x0), so the radix sort of the entries
// takes only the bits that box needs (the broad phase has checked that every
// bound is below 2^20 cells, so X Y Z < 2^63).
struct BinFrame {
    int x0, y0, z0;
    uint64_t X, Y;
    __device__ __forceinline__ uint64_t code(int x, int y, int z) const {
        return (uint64_t(z - z0) * Y + uint64_t(y - y0)) * X + uint64_t(x - x0);
    }
    __device__ __forceinline__ void decode(uint64_t c, int &x, int &y, int &z) const {
        x = int(c % X) + x0;
        c /= X;
        y = int(c % Y) + y0;
        z = int(c / Y) + z0;
    }
};

// One entry per covered bin; `ends` is the inclusive scan of the counts.
__global__ void k_emit_bins(uint32_t n, const int32_t *__restrict__ lohi, const uint64_t *__restrict__ count,
                            const uint64_t *__restrict__ ends, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ vals, BinFrame fr) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint64_t o = ends[v] - count[v];
    const int32_t *b = lohi + size_t(v) * 6;
    for (int z = b[2]; z <= b[5]; ++z)
        for (int y = b[1]; y <= b[4]; ++y)
            for (int x = b[0]; x <= b[3]; ++x) {
                keys[o] = fr.code(x, y, z);
                vals[o] = v;
                ++o;
            }
}

// local_continuum_pair (src/collision.rs:520-550): the offset of the rest
// positions rounds to one lattice step or less per axis and a monotone path
// of intact bonds walks it. The recursion of the reference tries the axes in
// every order; so does this loop over the six permutations.
__device__ bool local_continuum(const float *__restrict__ rest, const uint32_t *__restrict__ adj, uint32_t i,
                                uint32_t j) {
    float steps[3], frac[3];
    for (int a = 0; a < 3; ++a) {
        const float d = (rest[size_t(j) * 3 + a] - rest[size_t(i) * 3 + a]) / kCell;
        steps[a] = roundf(d);
        frac[a] = fabsf(d - steps[a]);
    }
    const float fmax_ = fmaxf(frac[0], fmaxf(frac[1], frac[2]));
    const float smax = fmaxf(fabsf(steps[0]), fmaxf(fabsf(steps[1]), fabsf(steps[2])));
    if (fmax_ > 1e-3f || smax > 1.0f) return false;
    constexpr int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        uint32_t node = i;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (node == kEmpty) break;
            const int a = perms[p][k];
            if (steps[a] == 0.0f) continue;
            node = adj[size_t(node) * 6 + 2 * a + (steps[a] > 0.0f ? 1 : 0)];
        }
        if (node == j) return true;
    }
    return false;
}

// For entry e of the bin-sorted list: the candidate pairs it forms with the
// later entries of its bin, counted (EMIT = false) or written (EMIT = true).
// Pairs are written as (i << nb) | j with nb bits per voxel index (the radix
// sort then takes 2 nb bits) and unpacked to (i << 32) | j after the sort.
// The counting pass leaves in kmask[e] which of the entry's next 63 bin
// mates it keeps (bit 63: a kept mate further on, so the writing pass walks
// the bin again), so the writing pass repeats no test.
constexpr uint64_t kMaskWalk = uint64_t(1) << 63;
template <bool EMIT>
__global__ void k_bin_pairs(uint64_t total, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                            const int32_t *__restrict__ lohi, const uint8_t *__restrict__ flags,
                            const float *__restrict__ rest, const uint32_t *__restrict__ adj,
                            uint64_t *__restrict__ count, const uint64_t *__restrict__ ends,
                            uint64_t *__restrict__ out, unsigned long long *__restrict__ candidates, BinFrame fr,
                            uint32_t nb, uint64_t *__restrict__ kmask) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long cand = 0;
    if (e < total) {
        const uint32_t v = vals[e];
        uint64_t o = EMIT ? ends[e] - count[e] : 0;
        const uint64_t mask = EMIT ? kmask[e] : 0;
        if (EMIT && !(mask & kMaskWalk)) {
            for (uint64_t b = mask; b; b &= b - 1) {
                const uint32_t u = vals[e + 1 + uint32_t(__ffsll((unsigned long long)b) - 1)];
                out[o++] = (uint64_t(min(u, v)) << nb) | max(u, v);
            }
        } else {
            const uint64_t key = keys[e];
            const bool kin_v = kinematic(flags[v]);
            int bx, by, bz;
            fr.decode(key, bx, by, bz);
            uint64_t kept = 0, bits = 0;
            const int32_t *lv = lohi + size_t(v) * 6;
            const int lv0 = lv[0], lv1 = lv[1], lv2 = lv[2];
            for (uint64_t f = e + 1; f < total && keys[f] == key; ++f) {
                const uint32_t u = vals[f];
                if (kin_v && kinematic(flags[u])) continue;
                // emit from the lowest corner of the two ranges' intersection only
                const int32_t *lu = lohi + size_t(u) * 6;
                if (max(lu[0], lv0) != bx || max(lu[1], lv1) != by || max(lu[2], lv2) != bz) continue;
                ++cand;
                const uint32_t i = min(u, v), j = max(u, v);
                if (local_continuum(rest, adj, i, j)) continue;
                if (EMIT) out[o++] = (uint64_t(i) << nb) | j;
                bits |= f - e - 1 < 63 ? uint64_t(1) << (f - e - 1) : kMaskWalk;
                ++kept;
            }
            if (!EMIT) {
                count[e] = kept;
                kmask[e] = bits;
            }
        }
    }
    if (!EMIT) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) cand += __shfl_xor(cand, off, 64);
        __shared__ unsigned long long red[4];
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cand;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t = 0;
            for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += red[w];
            candidates[blockIdx.x] = t;  // per workgroup; k_sum_reduce adds them
        }
    }
}
__global__ void k_unpack_pairs(uint32_t P, uint64_t *pairs, uint32_t nb) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= P) return;
    const uint64_t k = pairs[m];
    pairs[m] = ((k >> nb) << 32) | (k & ((uint64_t(1) << nb) - 1));
}

// ---- narrow phase: the reference's tests, f32 ---------------------------------

struct V3 {
    float x[3];
};
__device__ __forceinline__ V3 sub3(const V3 &a, const V3 &b) {
    return V3{{a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]}};
}
// glam Vec3::dot: (x*x' + y*y') + z*z'
__device__ __forceinline__ float dot3(const V3 &a, const V3 &b) {
    return a.x[0] * b.x[0] + a.x[1] * b.x[1] + a.x[2] * b.x[2];
}
// Component `axis` of v without a runtime-indexed array (which would live in memory).
__device__ __forceinline__ float comp(const V3 &v, int axis) {
    return axis == 0 ? v.x[0] : (axis == 1 ? v.x[1] : v.x[2]);
}
__device__ __forceinline__ V3 axis_normal(int axis, float sign) {
    return V3{{axis == 0 ? sign : 0.0f, axis == 1 ? sign : 0.0f, axis == 2 ? sign : 0.0f}};
}

// swept_normal (src/collision.rs:552-596): the axis, or -1.
__device__ int swept_normal(const V3 &si, const V3 &sj, const V3 &ei, const V3 &ej, float *sign_out) {
    const V3 d0 = sub3(sj, si), rel = sub3(sub3(ej, sj), sub3(ei, si));
    const float close = kCell - 1e-6f;
    if (fabsf(d0.x[0]) <= close && fabsf(d0.x[1]) <= close && fabsf(d0.x[2]) <= close) return -1;
    float enter = 0.0f, exit = 1.0f, sign = 1.0f;
    int hit = -1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (fabsf(rel.x[a]) < 1e-12f) {
            if (fabsf(d0.x[a]) > kCell) return -1;
            continue;
        }
        const float t0 = (-kCell - d0.x[a]) / rel.x[a];
        const float t1 = (kCell - d0.x[a]) / rel.x[a];
        const float near_t = fminf(t0, t1), far_t = fmaxf(t0, t1);
        if (near_t > enter || (hit < 0 && near_t >= enter)) {
            enter = near_t;
            hit = a;
            sign = (d0.x[a] + rel.x[a] * enter >= 0.0f) ? 1.0f : -1.0f;
        }
        exit = fminf(exit, far_t);
        if (enter > exit) return -1;
    }
    if (exit < 0.0f || enter > 1.0f) return -1;
    *sign_out = sign;
    return hit;
}

// overlap_normal (src/collision.rs:598-621): the axis, or -1.
__device__ int overlap_normal(const V3 &a, const V3 &b, const V3 &fallback, float *sign_out, float *depth) {
    const V3 d = sub3(b, a);
    const float o0 = kCell - fabsf(d.x[0]), o1 = kCell - fabsf(d.x[1]), o2 = kCell - fabsf(d.x[2]);
    if (fminf(o0, fminf(o1, o2)) <= 0.0f) return -1;
    const int axis = (o0 <= o1 && o0 <= o2) ? 0 : (o1 <= o2 ? 1 : 2);
    float src = 1.0f;
    const float da = comp(d, axis), fa = comp(fallback, axis);
    if (fabsf(da) > 1e-8f)
        src = da;
    else if (fabsf(fa) > 1e-8f)
        src = fa;
    *sign_out = signbit(src) ? -1.0f : 1.0f;  // f32::signum of a nonzero value
    *depth = axis == 0 ? o0 : (axis == 1 ? o1 : o2);
    return axis;
}

// touching_normal (src/collision.rs:441-460): the axis, or -1.
__device__ int touching_normal(const V3 &a, const V3 &b, float *sign_out) {
    const float slop = 2.0e-6f;
    const V3 d = sub3(b, a);
    const V3 g{{fabsf(d.x[0]) - kCell, fabsf(d.x[1]) - kCell, fabsf(d.x[2]) - kCell}};
    const int axis = (g.x[0] >= g.x[1] && g.x[0] >= g.x[2]) ? 0 : (g.x[1] >= g.x[2] ? 1 : 2);
    if (comp(g, axis) > slop) return -1;
#pragma unroll
    for (int o = 0; o < 3; ++o)
        if (o != axis && g.x[o] >= -slop) return -1;
    *sign_out = comp(d, axis) >= 0.0f ? 1.0f : -1.0f;
    return axis;
}

// inverse_mass (src/collision.rs:633-640)
__device__ __forceinline__ float inverse_mass(uint8_t f, float mass, int axis) {
    return (kinematic(f) || ((f >> (1 + axis)) & 1)) ? 0.0f : 1.0f / mass;
}

// The outcome of one pair evaluated on the current positions.
struct Outcome {
    uint32_t i = 0, j = 0;
    uint8_t code = 0;       // 0: no contact; else 0x80 | axis | 4 (normal negative) | 8 (new impact)
    bool sep = false;       // separate() ran with a positive denominator (src/collision.rs:650-652)
    bool movable_i = false, movable_j = false;  // not kinematic
    uint8_t real_i = 0, real_j = 0;  // components with a nonzero delta
    uint8_t plus0_i = 0, plus0_j = 0;  // components with a +0 delta that can meet a -0
    float di[3] = {0, 0, 0}, dj[3] = {0, 0, 0};
    // when sep: the normal axis, its sign, and the amounts separate() moves i and j by
    int8_t axis = 0;
    bool negative = false;
    float s_i = 0.0f, s_j = 0.0f;
};

__device__ __forceinline__ bool is_zero(float x) { return (__float_as_uint(x) & 0x7fffffffu) == 0; }

// True only when swept_normal (src/collision.rs:552-596) certainly returns
// None, decided without its divisions: the starts already overlap on every
// axis, or on some axis the pair starts apart and stays outside the slab
// |d| <= CELL for the whole step. Still-standing pairs (|rel| < 1e-12) return
// None in the reference right there; a receding pair has a slab exit time
// below 0; an approaching pair enters after t = 1 when its gap exceeds
// 1.002 |rel| (the margin covers the rounding of both sides, so the
// reference's rounded entry time is above 1 as well). Either way the
// reference returns None at that axis whatever the axes before it did.
__device__ __forceinline__ bool swept_surely_none(const V3 &d0, const V3 &rel) {
    const float close = kCell - 1e-6f;
    if (fabsf(d0.x[0]) <= close && fabsf(d0.x[1]) <= close && fabsf(d0.x[2]) <= close) return true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float d = d0.x[a], r = rel.x[a];
        if (fabsf(r) < 1e-12f) {
            if (fabsf(d) > kCell) return true;
        } else if (d > kCell) {
            if (r > 0.0f || (d - kCell) > 1.002f * -r) return true;
        } else if (d < -kCell) {
            if (r < 0.0f || (-kCell - d) > 1.002f * r) return true;
        }
    }
    return false;
}

// One pass body of project (src/collision.rs:55-109) for pair (i, j).
// The projection reads each voxel as one 32-byte record: position (3),
// start (3), mass, flags; two 16-byte loads from one cache line. The flags
// word carries the flag byte and, from bit kNegZeroShift, the components that
// are -0 on entry in the position or the previous position: a real move never
// gives -0 and a +0 only turns -0 into +0, so no other component ever needs
// the +0-over--0 store.
constexpr int kRec = 8;  // floats per voxel record
constexpr int kNegZeroShift = 8;
// The pass body on explicit positions pi, pj; the starts, masses and flags
// come from the voxel records (a0.w, a1 and b0.w, b1).
__device__ __forceinline__ Outcome evaluate_at(uint32_t i, uint32_t j, const V3 &pi, const float4 &a0,
                                               const float4 &a1, const V3 &pj, const float4 &b0, const float4 &b1) {
    Outcome o;
    o.i = i;
    o.j = j;
    const uint32_t wi_flags = __float_as_uint(a1.w), wj_flags = __float_as_uint(b1.w);
    const uint8_t fi = uint8_t(wi_flags), fj = uint8_t(wj_flags);
    // components that were -0 on entry (position or previous): the only ones a +0 can change
    const uint8_t nz_i = uint8_t((wi_flags >> kNegZeroShift) & 7u), nz_j = uint8_t((wj_flags >> kNegZeroShift) & 7u);
    o.movable_i = !kinematic(fi);
    o.movable_j = !kinematic(fj);
    const V3 si{{a0.w, a1.x, a1.y}};
    const V3 sj{{b0.w, b1.x, b1.y}};
    const V3 start_delta = sub3(sj, si);
    float sign = 1.0f, amount = 0.0f, depth = 0.0f;
    {
        // most pairs are apart: no contact from any of the three tests
        const V3 d = sub3(pj, pi);
        const bool no_overlap =
            fminf(kCell - fabsf(d.x[0]), fminf(kCell - fabsf(d.x[1]), kCell - fabsf(d.x[2]))) <= 0.0f;
        if (no_overlap && touching_normal(pi, pj, &sign) < 0 &&
            swept_surely_none(start_delta, sub3(sub3(pj, sj), sub3(pi, si))))
            return o;
    }
    bool contact = false, impact = false, do_sep = false;
    int axis = swept_normal(si, sj, pi, pj, &sign);
    if (axis >= 0) {
        const V3 nrm = axis_normal(axis, sign);
        const float separation = dot3(sub3(pj, pi), nrm);
        if (separation < kCell) {
            amount = kCell - separation + 1e-6f;
            do_sep = true;
        }
        contact = true;
        impact = dot3(start_delta, nrm) > kCell + 2e-6f;
    } else if ((axis = overlap_normal(pi, pj, start_delta, &sign, &depth)) >= 0) {
        amount = depth + 1e-6f;
        do_sep = true;
        contact = true;
        impact = dot3(start_delta, axis_normal(axis, sign)) > kCell + 2e-6f;
    } else if ((axis = touching_normal(pi, pj, &sign)) >= 0) {
        contact = true;
    }
    if (contact) o.code = uint8_t(0x80 | axis | (sign < 0.0f ? 4 : 0) | (impact ? 8 : 0));
    if (do_sep) {
        // separate (src/collision.rs:642-665)
        const float wi = inverse_mass(fi, a1.z, axis);
        const float wj = inverse_mass(fj, b1.z, axis);
        const float denom = wi + wj;
        if (denom > 0.0f) {
            o.sep = true;
            const float s_i = amount * wi / denom, s_j = amount * wj / denom;
            o.axis = int8_t(axis);
            o.negative = sign < 0.0f;
            o.s_i = s_i;
            o.s_j = s_j;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float nc = c == axis ? sign : 0.0f;
                o.di[c] = -nc * s_i;
                o.dj[c] = nc * s_j;
                if (!is_zero(o.di[c])) o.real_i |= uint8_t(1 << c);
                else if (__float_as_uint(o.di[c]) == 0u) o.plus0_i |= uint8_t((1 << c) & nz_i);
                if (!is_zero(o.dj[c])) o.real_j |= uint8_t(1 << c);
                else if (__float_as_uint(o.dj[c]) == 0u) o.plus0_j |= uint8_t((1 << c) & nz_j);
            }
        }
    }
    return o;
}
__device__ Outcome evaluate(uint32_t i, uint32_t j, const float *rec) {
    const float4 *r4 = reinterpret_cast<const float4 *>(rec);
    const float4 a0 = r4[2 * size_t(i)], a1 = r4[2 * size_t(i) + 1];
    const float4 b0 = r4[2 * size_t(j)], b1 = r4[2 * size_t(j) + 1];
    return evaluate_at(i, j, V3{{a0.x, a0.y, a0.z}}, a0, a1, V3{{b0.x, b0.y, b0.z}}, b0, b1);
}

// ---- projection ----------------------------------------------------------------

constexpr int kR = 2;                        // pairs per thread and window
constexpr uint32_t kWin = uint32_t(kGs) * kR;  // window slots: s = r * kGs + threadIdx.x

// One window slot. A slot holds a pair until the pair commits; a committed or
// empty slot takes the next pair of the list. Pairs are ordered by their index
// k in the sorted list, whatever slot they sit in. Flags are packed in one
// word to keep kR slots in registers.
enum : uint32_t {
    S_LIVE = 1u << 0, S_SEP = 1u << 2, S_MOV_I = 1u << 3, S_MOV_J = 1u << 4,
    S_REAL_I = 1u << 5, S_REAL_J = 1u << 6, S_HELD = 1u << 7, S_MARKED = 1u << 8,
};
constexpr int S_AXIS = 9;     // 2 bits: the normal axis
constexpr int S_PLUS0 = 11;   // 6 bits: +0 components of i (3) and j (3)
constexpr int S_CODE = 24;    // 8 bits: contact code
struct Slot {
    uint32_t k, m, i, j, f;  // k: position in the list of the launch; m: index in the sorted pairs
    float di, dj;  // the deltas along the normal axis (the others are +-0)
};

__device__ __forceinline__ void set_outcome(Slot &sl, const Outcome &o) {
    uint32_t f = sl.f & S_LIVE;
    if (o.sep) f |= S_SEP;
    if (o.movable_i) f |= S_MOV_I;
    if (o.movable_j) f |= S_MOV_J;
    if (o.real_i) f |= S_REAL_I;
    if (o.real_j) f |= S_REAL_J;
    f |= (uint32_t(o.plus0_i) | (uint32_t(o.plus0_j) << 3)) << S_PLUS0;
    f |= uint32_t(o.code) << S_CODE;
    uint32_t axis = 0;
    float di = 0.0f, dj = 0.0f;
    // only the normal axis can carry a nonzero delta
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (o.real_i & (1 << c)) {
            axis = uint32_t(c);
            di = o.di[c];
        }
        if (o.real_j & (1 << c)) {
            axis = uint32_t(c);
            dj = o.dj[c];
        }
    }
    sl.f = f | (axis << S_AXIS);
    sl.di = di;
    sl.dj = dj;
}

// Wait until this wave's global stores and atomics are performed, so that the
// barrier after it publishes them to the other waves.
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

// Window marks live in an LDS hash table keyed by voxel (open addressing,
// linear probing). A window holds kWin pairs, so at most 2 * kWin distinct
// voxels are inserted: the table is twice that, never full, and every probe
// sequence ends at an empty entry. hw[h]: lowest pair index of the window that
// moves the voxel; hb[h]: lowest held pair index that touches it (movable
// voxels only).
constexpr uint32_t kHash = 4 * kWin;
constexpr uint32_t kFreeKey = 0xFFFFFFFFu;  // voxel indices are below n <= 2^32 - 1
__device__ __forceinline__ uint32_t hash_slot(uint32_t v) { return (v * 2654435761u) >> 19; }  // 13 bits
static_assert(kHash == 8192, "hash_slot gives 13 bits");
__device__ __forceinline__ void mark(uint32_t *hkey, uint32_t *hval, uint32_t v, uint32_t k) {
    uint32_t h = hash_slot(v);
    for (;;) {
        const uint32_t was = atomicCAS(hkey + h, kFreeKey, v);
        if (was == kFreeKey || was == v) {
            atomicMin(hval + h, k);
            return;
        }
        h = (h + 1) & (kHash - 1);
    }
}
__device__ __forceinline__ uint32_t marked(const uint32_t *hkey, const uint32_t *hval, uint32_t v) {
    uint32_t h = hash_slot(v);
    for (;;) {
        const uint32_t key = hkey[h];
        if (key == v) return hval[h];
        if (key == kFreeKey) return kNoMark;
        h = (h + 1) & (kHash - 1);
    }
}

template <int STRIDE>
__device__ __forceinline__ void add_plus0(float *p, uint32_t v, uint32_t mask) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if ((mask & (1u << c)) && __float_as_uint(p[size_t(v) * STRIDE + c]) == 0x80000000u)
            p[size_t(v) * STRIDE + c] = 0.0f;
}
// The deferred +0 stores of one committed pair (mask: i in bits 0-2, j in 3-5).
__device__ __forceinline__ void apply_plus0(float *rec, float *prev, uint32_t i, uint32_t j, uint32_t mask) {
    add_plus0<kRec>(rec, i, mask & 7u);
    add_plus0<3>(prev, i, mask & 7u);
    add_plus0<kRec>(rec, j, mask >> 3);
    add_plus0<3>(prev, j, mask >> 3);
}

// Per-wave popcounts of NF flags for each of the kR slot rows, exchanged
// through LDS in one barrier. total[f] is the workgroup count of flag f;
// rank[r] is the position of slot (r, thread) among the flag-0 slots in slot
// order. The two LDS rows alternate between calls: a row is rewritten only
// after every thread has passed the next call's barrier.
template <int NF>
__device__ __forceinline__ void slot_counts(const bool (&f)[NF][kR], uint32_t (*cnt)[NF][kR][kGsWaves], int &row,
                                            uint32_t (&total)[NF], uint32_t (&rank)[kR]) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t b0[kR];
#pragma unroll
    for (int q = 0; q < NF; ++q)
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            const uint64_t b = __ballot(f[q][r]);
            if (q == 0) b0[r] = b;
            if (lane == 0) cnt[row][q][r][wave] = uint32_t(__popcll(b));
        }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        uint32_t t = 0;
        for (int r = 0; r < kR; ++r)
            for (int w = 0; w < kGsWaves; ++w) t += cnt[row][q][r][w];
        total[q] = t;
    }
    uint32_t before = 0;
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        uint32_t in_row = 0;
        for (int w = 0; w < kGsWaves; ++w) in_row += (uint32_t(w) < wave) ? cnt[row][0][r][w] : 0u;
        rank[r] = before + in_row + uint32_t(__popcll(b0[r] & ((uint64_t(1) << lane) - 1)));
        for (int w = 0; w < kGsWaves; ++w) before += cnt[row][0][r][w];
    }
    row ^= 1;
}

// Control words of one call, in device memory (the passes are separate
// launches enqueued without host round trips).
enum : int {
    C_DONE = 0, C_OK, C_CHANGED, C_BAD, C_PASSES, C_WINDOWS, C_RESTORE, C_REDONE, C_EVALS, C_CANDS, C_FULL, C_PREV,
    C_FIRST, C_WORDS
};

// Workgroup minimum, broadcast; the two LDS rows alternate as in slot_counts.
__device__ __forceinline__ uint32_t block_min(uint32_t v, uint32_t (*mins)[kGsWaves], int &row) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, uint32_t(__shfl_xor(int(v), off, 64)));
    if ((threadIdx.x & 63) == 0) mins[row][threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t m = mins[row][0];
    for (int w = 1; w < kGsWaves; ++w) m = min(m, mins[row][w]);
    row ^= 1;
    return m;
}

// One Gauss-Seidel pass of project (src/collision.rs:54-109) over the sorted
// pairs (list == nullptr) or over the candidate pairs list[0, *list_count),
// window by window, in one workgroup. A pair k is held when an earlier pair
// of the window moves one of its voxels (hw below k) or an earlier held pair
// touches one of its movable voxels (hb below k), with the closure under that
// rule cut after one round as described below, before anything commits. Four
// barriers per window. The pair keys of the next window are fetched into LDS
// while the current one runs. The +0-over--0 stores of a window's committed
// pairs are made at the start of the next window: pair tests do not depend on
// the sign of a zero, and the next real moves come after that window's
// barriers. Per committed pair: its contact (first one kept), its outcome
// (sep_last) and, per moved voxel, the first and last moving pair
// (wfirst: index, wlast: index + 1). ctl: the control words; the launch returns at once when
// the pass is over or settled.
__global__ __launch_bounds__(kGs) void k_project(const uint64_t *__restrict__ pairs, const uint32_t *__restrict__ list,
                                                 const uint32_t *__restrict__ list_count, uint32_t npairs, float *rec,
                                                 float *prev, uint8_t *code, uint8_t *sep_last, uint32_t *wfirst,
                                                 uint32_t *wlast, uint32_t *ctl, int from_first = 0) {
    if (ctl[C_DONE] | ctl[C_OK] | (list ? ctl[C_FULL] : 0u)) return;
    // from_first: the first pass, from its first separating pair on (k_precheck)
    const uint32_t first = (!list && from_first) ? min(ctl[C_FIRST], npairs) : 0u;
    const uint32_t count = list ? *list_count : npairs - first;
    __shared__ uint32_t hkey[kHash], hw[kHash], hb[kHash];
    __shared__ uint64_t pre[2][kWin];
    __shared__ uint32_t prem[2][kWin];
    __shared__ uint32_t cnt2[2][2][kR][kGsWaves];
    __shared__ uint32_t mins[2][kGsWaves];
    const uint32_t t = threadIdx.x;
    int row2 = 0, row_min = 0;
    for (uint32_t x = t; x < kHash; x += kGs) {
        hkey[x] = kFreeKey;
        hw[x] = kNoMark;
        hb[x] = kNoMark;
    }
    __syncthreads();
    uint32_t windows = 0, buf = 0;
    Slot sl[kR];
    uint32_t fx_i[kR], fx_j[kR], fx_m[kR];  // deferred +0 stores: voxels and masks (i: bits 0-2, j: 3-5)
#pragma unroll
    for (int r = 0; r < kR; ++r) fx_m[r] = 0;
    {
        bool changed = false;
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            const uint32_t s = uint32_t(r) * kGs + t;
            sl[r].k = s;
            sl[r].f = s < count ? S_LIVE : 0u;
            if (s < count) {
                sl[r].m = list ? list[s] : first + s;
                const uint64_t key = pairs[sl[r].m];
                sl[r].i = uint32_t(key >> 32);
                sl[r].j = uint32_t(key);
            }
        }
        uint32_t next = min(kWin, count);
        uint32_t live_total = next;
        while (live_total > 0) {
            ++windows;
            // keys of the pairs the next window may take: pairs[next + x]
            uint64_t ahead[kR];
            uint32_t ahead_m[kR];
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                const uint32_t x = uint32_t(r) * kGs + t;
                ahead_m[r] = 0;
                ahead[r] = 0;
                if (x < count - next) {
                    ahead_m[r] = list ? list[next + x] : first + next + x;
                    ahead[r] = pairs[ahead_m[r]];
                }
            }
#pragma unroll
            for (int r = 0; r < kR; ++r)
                if (fx_m[r]) {
                    apply_plus0(rec, prev, fx_i[r], fx_j[r], fx_m[r]);
                    fx_m[r] = 0;
                }
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                Slot &q = sl[r];
                if (q.f & S_LIVE) set_outcome(q, evaluate(q.i, q.j, rec));
                if (q.f & S_REAL_I) mark(hkey, hw, q.i, q.k);
                if (q.f & S_REAL_J) mark(hkey, hw, q.j, q.k);
            }
            __syncthreads();
            // held: an earlier pair of the window moves one of its voxels ...
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                Slot &q = sl[r];
                if ((q.f & S_LIVE) && (marked(hkey, hw, q.i) < q.k || marked(hkey, hw, q.j) < q.k)) {
                    q.f |= S_HELD;
                    if (q.f & S_MOV_I) mark(hkey, hb, q.i, q.k);
                    if (q.f & S_MOV_J) mark(hkey, hb, q.j, q.k);
                }
            }
            __syncthreads();
            // ... or an earlier held pair touches one of its movable voxels. One
            // round of that closure; if it holds new pairs, the window is cut at the
            // smallest of them k*: pairs below k* are final (marks only hold later
            // pairs), so holding every pair from k* on closes the held set.
            uint32_t kmin = kNoMark;
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                Slot &q = sl[r];
                const bool newly = (q.f & S_LIVE) && !(q.f & S_HELD) &&
                                   (marked(hkey, hb, q.i) < q.k || marked(hkey, hb, q.j) < q.k);
                if (newly) kmin = min(kmin, q.k);
            }
            const uint32_t kstar = block_min(kmin, mins, row_min);
            if (kstar != kNoMark) {
#pragma unroll
                for (int r = 0; r < kR; ++r)
                    if ((sl[r].f & S_LIVE) && sl[r].k >= kstar) sl[r].f |= S_HELD;
            }
            // every table read of the window is done: clear it for the next one
            for (uint32_t x = t; x < kHash; x += kGs) {
                hkey[x] = kFreeKey;
                hw[x] = kNoMark;
                hb[x] = kNoMark;
            }
            bool fl2[2][kR];
            uint32_t rk[kR];
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                Slot &q = sl[r];
                const bool live = q.f & S_LIVE;
                const bool commit = live && !(q.f & S_HELD);
                const bool sep = commit && (q.f & S_SEP);
                if (sep) {
                    const size_t a = (q.f >> S_AXIS) & 3u;
                    if (q.f & S_REAL_I) {
                        rec[size_t(q.i) * kRec + a] += q.di;
                        prev[size_t(q.i) * 3 + a] += q.di;
                    }
                    if (q.f & S_REAL_J) {
                        rec[size_t(q.j) * kRec + a] += q.dj;
                        prev[size_t(q.j) * 3 + a] += q.dj;
                    }
                    fx_i[r] = q.i;
                    fx_j[r] = q.j;
                    fx_m[r] = (q.f >> S_PLUS0) & 63u;
                    if (q.f & S_REAL_I) {
                        atomicMin(wfirst + q.i, q.m);
                        atomicMax(wlast + q.i, q.m + 1);
                    }
                    if (q.f & S_REAL_J) {
                        atomicMin(wfirst + q.j, q.m);
                        atomicMax(wlast + q.j, q.m + 1);
                    }
                }
                // the contact of the first pass that finds one (contacts are deduplicated per pair)
                const uint8_t c = uint8_t(q.f >> S_CODE);
                if (commit && c && code[q.m] == 0) code[q.m] = c;
                if (commit) sep_last[q.m] = sep ? 1 : 0;
                fl2[0][r] = !live || commit;  // the slot takes a new pair
                fl2[1][r] = sep;
            }
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                pre[buf][uint32_t(r) * kGs + t] = ahead[r];
                prem[buf][uint32_t(r) * kGs + t] = ahead_m[r];
            }
            wait_stores();  // the moves are performed before the next window reads positions
            uint32_t tot2[2];
            slot_counts<2>(fl2, cnt2, row2, tot2, rk);
            if (tot2[1] > 0) changed = true;
            const uint32_t added = min(tot2[0], count - next);
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                Slot &q = sl[r];
                if (fl2[0][r]) {
                    q.k = next + rk[r];
                    q.f = rk[r] < added ? S_LIVE : 0u;
                    if (rk[r] < added) {
                        q.m = prem[buf][rk[r]];
                        const uint64_t key = pre[buf][rk[r]];
                        q.i = uint32_t(key >> 32);
                        q.j = uint32_t(key);
                    }
                } else {
                    q.f = S_LIVE;  // held: evaluated again on the next window
                }
            }
            buf ^= 1;
            next += added;
            live_total = (kWin - tot2[0]) + added;
        }
        if (changed && t == 0) ctl[C_CHANGED] = 1;
    }
#pragma unroll
    for (int r = 0; r < kR; ++r)
        if (fx_m[r]) apply_plus0(rec, prev, fx_i[r], fx_j[r], fx_m[r]);
    if (t == 0) {
        ctl[C_WINDOWS] += windows;
        ctl[C_EVALS] += count;
    }
}

// ---- passes after the first ------------------------------------------------------
//
// A pair whose two voxels have not moved since its last evaluation, and whose
// last evaluation did not separate, would give the same outcome again: no
// move, no new contact (the first one is kept), no `changed`. Only the other
// pairs need evaluating. A pass first takes as candidates the pairs that
// separated last time or touch a voxel moved in the previous pass (moves
// before a pair's own slot included: a superset). Moves inside the pass can
// make more pairs due; after the candidate pass, every skipped pair is checked
// against the first move of each of its voxels in this pass (wfirst below its
// index means it was due). If any was, the pass is undone from the snapshot
// and run again with those pairs added; after kAttempts it runs over every
// pair. The result is the sequential sweep's either way.
constexpr int kAttempts = 3;

__global__ void k_pass_begin(uint32_t *ctl) {
    if (ctl[C_DONE]) return;
    if (!ctl[C_CHANGED]) {  // the last pass moved nothing (src/collision.rs:105-107)
        ctl[C_DONE] = 1;
        return;
    }
    ctl[C_PASSES] += 1;
    ctl[C_OK] = 0;
    ctl[C_CHANGED] = 0;
    ctl[C_BAD] = 0;
    ctl[C_RESTORE] = 0;
    ctl[C_CANDS] = 0;
    ctl[C_FULL] = 0;
}

// Candidates past half the pairs: the pass goes over every pair at once, as
// attempts would cost more than they skip (k_settle applies the same rule,
// and gives up on attempts whose failures keep adding many pairs).
__global__ void k_choose(uint32_t P, uint32_t *ctl) {
    if (!ctl[C_DONE] && ctl[C_CANDS] > P / 2) ctl[C_FULL] = 1;
    ctl[C_PREV] = ctl[C_CANDS];
}

// Per voxel: the previous pass's moves, fresh move records, the snapshot.
__global__ void k_pass_voxels(uint32_t n, const uint32_t *ctl, uint32_t *moved, uint32_t *wfirst, uint32_t *wlast,
                              const float *__restrict__ rec, float *__restrict__ pos_s,
                              const float *__restrict__ prev, float *__restrict__ prev_s) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n || ctl[C_DONE]) return;
    moved[v] = wlast[v] != 0 ? 1u : 0u;
    wfirst[v] = kNoMark;
    wlast[v] = 0;
    for (int c = 0; c < 3; ++c) {
        pos_s[size_t(v) * 3 + c] = rec[size_t(v) * kRec + c];
        prev_s[size_t(v) * 3 + c] = prev[size_t(v) * 3 + c];
    }
}

// Per pair: the candidate flag and the snapshot of its contact and last outcome.
__global__ void k_pass_pairs(uint32_t P, uint32_t *ctl, const uint64_t *__restrict__ pairs,
                             const uint32_t *__restrict__ moved, const uint8_t *__restrict__ sep_last,
                             uint8_t *__restrict__ cflag, const uint8_t *__restrict__ code,
                             uint8_t *__restrict__ code_s, uint8_t *__restrict__ sep_s) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl[C_DONE]) return;
    bool due = false;
    if (m < P) {
        const uint64_t key = pairs[m];
        due = sep_last[m] || moved[uint32_t(key >> 32)] || moved[uint32_t(key)];
        cflag[m] = due ? 1 : 0;
        code_s[m] = code[m];
        sep_s[m] = sep_last[m];
    }
    const int c = __syncthreads_count(due);
    if (threadIdx.x == 0 && c) atomicAdd(ctl + C_CANDS, uint32_t(c));
}

// After a candidate pass: a skipped pair with a voxel moved before its slot was due.
__global__ void k_verify(uint32_t P, uint32_t *ctl, const uint64_t *__restrict__ pairs, uint8_t *cflag,
                         const uint32_t *__restrict__ wfirst) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl[C_DONE] || ctl[C_OK] || ctl[C_FULL]) return;
    bool due = false;
    if (m < P && !cflag[m]) {
        const uint64_t key = pairs[m];
        if (wfirst[uint32_t(key >> 32)] < m || wfirst[uint32_t(key)] < m) {
            cflag[m] = 1;
            due = true;
        }
    }
    const int c = __syncthreads_count(due);
    if (threadIdx.x == 0 && c) {
        atomicOr(ctl + C_BAD, 1u);
        atomicAdd(ctl + C_CANDS, uint32_t(c));
    }
}

__global__ void k_settle(uint32_t P, uint32_t *ctl) {
    if (ctl[C_DONE] || ctl[C_OK] || ctl[C_FULL]) return;
    if (!ctl[C_BAD]) {
        ctl[C_OK] = 1;
    } else {
        ctl[C_BAD] = 0;
        ctl[C_RESTORE] = 1;
        ctl[C_CHANGED] = 0;
        ctl[C_REDONE] += 1;
        // a failed attempt that adds a quarter more pairs means moves cascade
        // through the pass: attempts would keep failing
        if (ctl[C_CANDS] > P / 2 || ctl[C_CANDS] - ctl[C_PREV] >= ctl[C_PREV] / 4) ctl[C_FULL] = 1;
        ctl[C_PREV] = ctl[C_CANDS];
    }
}

__global__ void k_restore_voxels(uint32_t n, const uint32_t *ctl, float *__restrict__ rec,
                                 const float *__restrict__ pos_s, float *__restrict__ prev,
                                 const float *__restrict__ prev_s, uint32_t *wfirst, uint32_t *wlast) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n || !ctl[C_RESTORE]) return;
    for (int c = 0; c < 3; ++c) {
        rec[size_t(v) * kRec + c] = pos_s[size_t(v) * 3 + c];
        prev[size_t(v) * 3 + c] = prev_s[size_t(v) * 3 + c];
    }
    wfirst[v] = kNoMark;
    wlast[v] = 0;
}

__global__ void k_restore_pairs(uint32_t P, const uint32_t *ctl, uint8_t *__restrict__ code,
                                const uint8_t *__restrict__ code_s, uint8_t *__restrict__ sep_last,
                                const uint8_t *__restrict__ sep_s) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= P || !ctl[C_RESTORE]) return;
    code[m] = code_s[m];
    sep_last[m] = sep_s[m];
}

__global__ void k_restore_done(uint32_t *ctl) { ctl[C_RESTORE] = 0; }

// In the first pass, the pairs before the first one that separates move
// nothing, so each of their outcomes is its outcome on the starting
// positions. k_precheck evaluates every pair on the starting positions at
// once, keeps its contact code in code_s and finds that first separating pair
// (C_FIRST); k_first_pairs records the contacts of the pairs before it and
// k_first_done counts them, and the first window pass starts there. A pass in
// which no pair separates (C_FIRST = none) moves nothing and leaves `changed`
// false: the projection ends with it (src/collision.rs:105-107). code_s is
// the later passes' snapshot, written before they read it.
__global__ void k_precheck(uint32_t P, const uint64_t *__restrict__ pairs, const float *rec, uint8_t *__restrict__ code_s,
                           uint32_t *ctl) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t first = kNoMark;
    if (m < P) {
        const uint64_t key = pairs[m];
        const Outcome o = evaluate(uint32_t(key >> 32), uint32_t(key), rec);
        code_s[m] = o.code;
        if (o.sep) first = m;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) first = min(first, uint32_t(__shfl_xor(int(first), off, 64)));
    if ((threadIdx.x & 63) == 0 && first != kNoMark) atomicMin(ctl + C_FIRST, first);
}
__global__ void k_first_pairs(uint32_t P, const uint32_t *ctl, const uint8_t *__restrict__ code_s,
                              uint8_t *__restrict__ code) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= P || m >= ctl[C_FIRST]) return;
    code[m] = code_s[m];
}
__global__ void k_first_done(uint32_t P, uint32_t *ctl) {
    const uint32_t first = min(ctl[C_FIRST], P);
    ctl[C_EVALS] += first;
    if (first == P) ctl[C_DONE] = 1;
}

__global__ void k_ctl_init(uint32_t *ctl) {
    for (int w = 0; w < C_WORDS; ++w) ctl[w] = 0;
    ctl[C_PASSES] = 1;
    ctl[C_FIRST] = kNoMark;
}

// Voxel records for the projection, and the positions back out of them.
__global__ void k_pack(uint32_t n, const float *__restrict__ pos, const float *__restrict__ start,
                       const float *__restrict__ mass, const uint8_t *__restrict__ flags, float4 *__restrict__ rec) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const size_t b = size_t(v) * 3;
    rec[2 * size_t(v)] = make_float4(pos[b], pos[b + 1], pos[b + 2], start[b]);
    uint32_t nz = 0;
    for (int c = 0; c < 3; ++c)
        if (__float_as_uint(pos[b + c]) == 0x80000000u || __float_as_uint(start[b + c]) == 0x80000000u) nz |= 1u << c;
    rec[2 * size_t(v) + 1] = make_float4(start[b + 1], start[b + 2], kinematic(flags[v]) ? 1.0f : mass[v],
                                         __uint_as_float(uint32_t(flags[v]) | (nz << kNegZeroShift)));
}
__global__ void k_unpack(uint32_t n, const float *__restrict__ rec, float *__restrict__ pos) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    for (int c = 0; c < 3; ++c) pos[size_t(v) * 3 + c] = rec[size_t(v) * kRec + c];
}

// ---- contacts out ----------------------------------------------------------------

__global__ void k_contact_flags(uint32_t npairs, const uint8_t *__restrict__ code, uint64_t *__restrict__ flag) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < npairs) flag[k] = code[k] ? 1 : 0;
}

__global__ void k_contact_scatter(uint32_t npairs, const uint64_t *__restrict__ pairs,
                                  const uint8_t *__restrict__ code, const uint64_t *__restrict__ ends,
                                  vx_contact *__restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs || !code[k]) return;
    const uint8_t c = code[k];
    vx_contact r;
    r.i = uint32_t(pairs[k] >> 32);
    r.j = uint32_t(pairs[k]);
    for (int a = 0; a < 3; ++a) r.normal[a] = 0.0f;
    r.normal[c & 3] = (c & 4) ? -1.0f : 1.0f;
    r.new_impact = (c & 8) ? 1u : 0u;
    out[ends[k] - 1] = r;
}

// ---- passes by relaxation --------------------------------------------------------
//
// A pass's result is the unique solution of its own recurrence: pair k reads
// the positions its two voxels have after the moves of the earlier pairs that
// touch them (in pair order), and its outcome is the pass body on those
// positions. Evaluating every pair on the positions that its earlier pairs'
// current outcomes give, and repeating until no outcome changes, reaches that
// solution: a fixed point satisfies the recurrence pair by pair, in order. It
// takes about as many rounds as the longest chain of moves that feed each
// other (6-7 per pass on the 64^3 contact box), where the window walker needs
// one sequential window per conflict (9,329 there).
//
// State, all in device memory:
//  * per pair, its current outcome: rflag (separates, normal axis and sign,
//    contact code) and rsi, rsj (the amounts separate() moves i and j by);
//  * per voxel, its movers: the pairs with a separating outcome that touch it,
//    ascending, in mlist[moff[v], moff[v] + mcnt[v]) (room for every pair of
//    the voxel, so the list never overflows; bit 31 of an entry: the voxel is the
//    pair's i);
//  * per voxel, the pairs that touch it, ascending: jlist[joff, jend) holds
//    the pairs (i, v), and [pstart, pend) is the range of pairs (v, j) in the
//    sorted list (they follow every (i, v) since i < v).
// A round evaluates the dirty pairs on their before-states and records the
// outcomes that changed and the voxels they touch (k_rx_eval), then rebuilds
// those voxels' movers, marking dirty every pair after the smallest changed
// one (k_rx_rebuild). The dirty list may hold a pair twice (marked from both
// voxels); both evaluations give the same outcome. The
// positions stay as they were at the pass's start until the fixed point;
// then k_rx_commit_voxels applies each voxel's moves in pair order (every
// component, as separate() adds them, signed zeros included) and
// k_rx_commit_pairs keeps the first contacts and the outcomes the next pass
// starts from. A pass that overflows a list, or runs past the call's round
// budget or its own evaluation budget, is left to the window walker from its
// start, with the passes after it.
constexpr uint32_t kRxRounds = 1536;      // rounds per call before the walker takes the passes left (default)
constexpr uint64_t kRxEvalBudget = 16;    // pair evaluations per pass, in units of the pair count
constexpr uint32_t kRxChunk = 8;          // rounds enqueued between two host checks
constexpr uint32_t kRxMaxPairs = 1u << 30;
// control words: dirty-list and changed-voxel counts by round parity, overflow,
// rounds with dirty pairs (the evaluations, u64, follow at word 8)
enum : int { RX_CNT0 = 0, RX_CNT1, RX_VCNT0, RX_VCNT1, RX_OVER, RX_ROUNDS, RX_WORDS };
constexpr uint32_t RF_SEP = 1u, RF_NEG = 8u;  // rflag: bit 0 separates, bits 1-2 axis, bit 3 negative, 8-15 code
constexpr int kRxBlocks = 2048;           // grid of the list kernels (grid-stride loops), a pass's first round
constexpr int kRxBlocksLater = 512;       // at most, later rounds

__device__ __forceinline__ uint32_t rf_axis(uint32_t f) { return (f >> 1) & 3u; }

// The move separate() (src/collision.rs:642-665) gives one voxel of a pair:
// i moves by -n * s_i and j by +n * s_j, every component added.
__device__ __forceinline__ void add_move(float p[3], uint32_t f, float s, bool side_i) {
    const uint32_t axis = rf_axis(f);
    const float sign = (f & RF_NEG) ? -1.0f : 1.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float nc = uint32_t(c) == axis ? sign : 0.0f;
        p[c] += side_i ? -nc * s : nc * s;
    }
}

// Position of voxel v just before pair k: its position at the start of the
// pass plus the moves of its movers below k, in pair order.
__device__ __forceinline__ void before_state(float p[3], uint32_t v, uint32_t k, const uint32_t *__restrict__ mcnt,
                                             const uint32_t *__restrict__ moff, const uint32_t *__restrict__ mlist,
                                             const uint32_t *__restrict__ rflag, const float *__restrict__ rsi,
                                             const float *__restrict__ rsj) {
    const uint32_t c = mcnt[v];
    const uint32_t *ml = mlist + moff[v];
    for (uint32_t q = 0; q < c; ++q) {
        const uint32_t e = ml[q];
        const uint32_t m = e & 0x7FFFFFFFu;
        if (m >= k) break;
        const bool side_i = e >> 31;
        add_move(p, rflag[m], side_i ? rsi[m] : rsj[m], side_i);
    }
}

// Reserve `take` consecutive entries of a list per lane; every lane of the
// wave calls this together. Returns the lane's first entry, or kNoMark (and
// sets *over) when the list's capacity would be exceeded.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t take, uint32_t *count, uint32_t cap, uint32_t *over) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t x = take;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= uint32_t(d)) x += y;
    }
    const uint32_t total = __shfl(x, 63, 64);
    uint32_t base = 0;
    if (lane == 63 && total) base = atomicAdd(count, total);
    base = __shfl(base, 63, 64);
    const uint32_t first = base + x - take;
    if (take && uint64_t(first) + take > cap) {
        *over = 1;
        return kNoMark;
    }
    return first;
}

// Pairs per voxel. jkey[m] = j of pair m; sorted stably by j with the pair
// indices as values, the pairs (i, v) of each voxel v come out in ascending
// order: jlist[joff[v], jend[v]). The pairs (v, j) are the range
// [pstart[v], pend[v]) of the sorted list.
__global__ void k_rx_jkeys(uint32_t P, const uint64_t *__restrict__ pairs, uint32_t *jkey) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < P) jkey[m] = uint32_t(pairs[m]);
}
__global__ void k_rx_bounds(uint32_t P, const uint64_t *__restrict__ pairs, const uint32_t *__restrict__ jkey,
                            uint32_t *joff, uint32_t *jend, uint32_t *pstart, uint32_t *pend) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    const uint32_t j = jkey[t];
    if (t == 0 || jkey[t - 1] != j) joff[j] = t;
    if (t == P - 1 || jkey[t + 1] != j) jend[j] = t + 1;
    const uint32_t i = uint32_t(pairs[t] >> 32);
    if (t == 0 || uint32_t(pairs[t - 1] >> 32) != i) pstart[i] = t;
    if (t == P - 1 || uint32_t(pairs[t + 1] >> 32) != i) pend[i] = t + 1;
}

__global__ void k_rx_degrees(uint32_t n, const uint32_t *__restrict__ joff, const uint32_t *__restrict__ jend,
                             const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ pend, uint32_t *deg) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) deg[v] = (jend[v] - joff[v]) + (pend[v] - pstart[v]);
}
// The relaxation's state at the start of a call: no movers, no changed
// voxel, no stamps, no pair ranges, every pair without an outcome.
__global__ void k_rx_clear(uint32_t n, uint32_t P, uint32_t *mcnt, uint32_t *vchg, uint32_t *vstamp, uint32_t *joff,
                           uint32_t *jend, uint32_t *pstart, uint32_t *pend, uint32_t *rflag, float *rsi, float *rsj,
                           uint32_t *rx) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        mcnt[t] = 0;
        vchg[t] = kNoMark;
        vstamp[t] = 0;
        joff[t] = 0;
        jend[t] = 0;
        pstart[t] = 0;
        pend[t] = 0;
    }
    if (t < P) {
        rflag[t] = 0;
        rsi[t] = 0.0f;
        rsj[t] = 0.0f;
    }
    if (t < 16) rx[t] = 0;
}
// First pass: every pair is dirty, nothing moves yet.
__global__ void k_rx_init_first(uint32_t P, uint32_t *dlist, uint32_t *rx, uint32_t cur) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < P) dlist[m] = m;
    if (m == 0) {
        rx[RX_CNT0 + cur] = P;
        rx[RX_OVER] = 0;
    }
}
// Later passes. A pair whose voxels no separating pair touched in the previous
// pass, and which did not separate there itself, would repeat its outcome
// (the one it holds). The others are the pairs that touch a voxel with
// movers: they are listed dirty (twice when both voxels have movers) and
// start from no outcome, and the movers are cleared.
__global__ __launch_bounds__(256) void k_rx_init_pass(uint32_t n, uint32_t *mcnt, const uint32_t *__restrict__ joff,
                                                      const uint32_t *__restrict__ jend,
                                                      const uint32_t *__restrict__ jlist,
                                                      const uint32_t *__restrict__ pstart,
                                                      const uint32_t *__restrict__ pend, uint32_t *rflag, float *rsi,
                                                      float *rsj, uint32_t *dlist, uint32_t *rx, uint32_t cur,
                                                      uint32_t cap) {
    // each wave takes 64 voxels and walks the pairs of those with movers, two
    // at a time, 32 lanes per voxel
    const uint32_t lane = threadIdx.x & 63, half = lane >> 5, sub = lane & 31;
    const uint32_t v0 = (blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * 64;
    uint64_t todo = __ballot(v0 + lane < n && mcnt[v0 + lane] != 0);
    while (todo) {
        const int a = __ffsll((unsigned long long)todo) - 1;
        todo &= todo - 1;
        int b = -1;
        if (todo) {
            b = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
        }
        const int mine = half ? b : a;
        const bool act = mine >= 0;
        const uint32_t v = v0 + uint32_t(act ? mine : 0);
        uint32_t j0 = 0, nj = 0, r0 = 0, nr = 0;
        if (act) {
            j0 = joff[v];
            nj = jend[v] - j0;
            r0 = pstart[v];
            nr = pend[v] - r0;
        }
        const uint32_t total = nj + nr;
        uint32_t iters = (total + 31) / 32;
        iters = max(iters, uint32_t(__shfl_xor(int(iters), 32, 64)));
        for (uint32_t it = 0; it < iters; ++it) {
            const uint32_t t = it * 32 + sub;
            const bool live = t < total;
            uint32_t m = 0;
            if (live) {
                m = t < nj ? jlist[j0 + t] : r0 + (t - nj);
                rflag[m] = 0;
                rsi[m] = 0.0f;
                rsj[m] = 0.0f;
            }
            const uint32_t pos = wave_reserve(live ? 1u : 0u, rx + RX_CNT0 + cur, cap, rx + RX_OVER);
            if (live && pos != kNoMark) dlist[pos] = m;
        }
        if (act && sub == 0) mcnt[v] = 0;
    }
}
__global__ void k_rx_init_words(uint32_t *rx, uint32_t cur) {
    rx[RX_CNT0 + cur] = 0;
    rx[RX_OVER] = 0;
}

__device__ __forceinline__ void pack_rx(const Outcome &o, uint32_t &f, float &si, float &sj) {
    f = uint32_t(o.code) << 8;
    si = 0.0f;
    sj = 0.0f;
    if (o.sep) {
        f |= RF_SEP | (uint32_t(o.axis) << 1) | (o.negative ? RF_NEG : 0u);
        si = o.s_i;
        sj = o.s_j;
    }
}

// Evaluate the dirty pairs of list `cur` on their before-states and record
// the outcomes that changed: a changed move marks both voxels (the smallest
// changed pair per voxel in vchg; each voxel listed once per round by its
// stamp); a changed contact code alone changes no position. A pair may read a
// mover that another thread of this round is rewriting (a torn or newer
// value): that mover's move changed, so the rebuild lists the pair again for
// the next round, and the round that ends the relaxation changes no move, so
// every pair's last evaluation read settled values.
__global__ __launch_bounds__(256) void k_rx_eval(uint32_t cur, uint32_t stamp, const uint32_t *__restrict__ dlist,
                                                 uint32_t *rx, uint32_t cap, const uint64_t *__restrict__ pairs,
                                                 const float *__restrict__ rec, const uint32_t *__restrict__ mcnt,
                                                 const uint32_t *__restrict__ moff, const uint32_t *__restrict__ mlist,
                                                 uint32_t *rflag, float *rsi, float *rsj, uint32_t *vchg,
                                                 uint32_t *vstamp, uint32_t *vlist, uint32_t n,
                                                 unsigned long long *evals) {
    if (rx[RX_OVER]) return;
    const uint32_t count = min(rx[RX_CNT0 + cur], cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the lists the previous round's rebuild consumed take this round's rebuild's output
        rx[RX_CNT0 + (cur ^ 1u)] = 0;
        rx[RX_VCNT0 + (cur ^ 1u)] = 0;
        if (count) {
            atomicAdd(evals, (unsigned long long)count);
            rx[RX_ROUNDS] += 1;
        }
    }
    const float4 *r4 = reinterpret_cast<const float4 *>(rec);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t x = base + threadIdx.x;
        uint32_t add_i = 0, add_j = 0, i = 0, j = 0;
        if (x < count) {
            const uint32_t k = dlist[x];
            const uint64_t key = pairs[k];
            i = uint32_t(key >> 32);
            j = uint32_t(key);
            const float4 a0 = r4[2 * size_t(i)], a1 = r4[2 * size_t(i) + 1];
            const float4 b0 = r4[2 * size_t(j)], b1 = r4[2 * size_t(j) + 1];
            float pi[3] = {a0.x, a0.y, a0.z}, pj[3] = {b0.x, b0.y, b0.z};
            before_state(pi, i, k, mcnt, moff, mlist, rflag, rsi, rsj);
            before_state(pj, j, k, mcnt, moff, mlist, rflag, rsi, rsj);
            const Outcome o = evaluate_at(i, j, V3{{pi[0], pi[1], pi[2]}}, a0, a1, V3{{pj[0], pj[1], pj[2]}}, b0, b1);
            uint32_t f;
            float si, sj;
            pack_rx(o, f, si, sj);
            const uint32_t of = rflag[k];
            const bool moves = ((f ^ of) & 0xFu) || __float_as_uint(si) != __float_as_uint(rsi[k]) ||
                               __float_as_uint(sj) != __float_as_uint(rsj[k]);
            if (moves || f != of) {
                rflag[k] = f;
                rsi[k] = si;
                rsj[k] = sj;
            }
            if (moves) {
                atomicMin(vchg + i, k);
                atomicMin(vchg + j, k);
                add_i = atomicExch(vstamp + i, stamp) != stamp ? 1u : 0u;
                add_j = atomicExch(vstamp + j, stamp) != stamp ? 1u : 0u;
            }
        }
        const uint32_t pos = wave_reserve(add_i + add_j, rx + RX_VCNT0 + cur, n, rx + RX_OVER);
        if (pos != kNoMark) {
            uint32_t q = pos;
            if (add_i) vlist[q++] = i;
            if (add_j) vlist[q] = j;
        }
    }
}

// Rebuild the movers of the voxels whose moves changed, and list as dirty the
// pairs after the smallest changed one that touch them. Two voxels per wave,
// 32 lanes each walking the voxel's pairs together in pair order (the sorted
// (i, v) list, then the (v, j) range); the movers are placed by ballot.
__global__ __launch_bounds__(256) void k_rx_rebuild(uint32_t cur, uint32_t *rx, uint32_t cap,
                                                    const uint32_t *__restrict__ vlist, const uint32_t *__restrict__ joff,
                                                    const uint32_t *__restrict__ jend,
                                                    const uint32_t *__restrict__ jlist,
                                                    const uint32_t *__restrict__ pstart,
                                                    const uint32_t *__restrict__ pend,
                                                    const uint32_t *__restrict__ rflag, uint32_t *mcnt,
                                                    const uint32_t *__restrict__ moff, uint32_t *mlist, uint32_t *vchg,
                                                    uint32_t *dnext) {
    if (rx[RX_OVER]) return;
    const uint32_t count = rx[RX_VCNT0 + cur];
    const uint32_t lane = threadIdx.x & 63, half = lane >> 5, sub = lane & 31;
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    const uint32_t wave_id = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    for (uint32_t base = wave_id * 2; base < count; base += waves * 2) {
        const uint32_t x = base + half;
        const bool act = x < count;
        uint32_t v = 0, from = kNoMark, j0 = 0, nj = 0, r0 = 0, nr = 0, mo = 0;
        if (act) {
            v = vlist[x];
            from = vchg[v];
            mo = moff[v];
            j0 = joff[v];
            nj = jend[v] - j0;
            r0 = pstart[v];
            nr = pend[v] - r0;
        }
        const uint32_t total = nj + nr;
        uint32_t iters = (total + 31) / 32;
        iters = max(iters, uint32_t(__shfl_xor(int(iters), 32, 64)));
        uint32_t c = 0;  // movers placed so far (the same in every lane of the half)
        for (uint32_t it = 0; it < iters; ++it) {
            const uint32_t t = it * 32 + sub;
            const bool live = act && t < total;
            uint32_t m = 0;
            bool sep = false, dirty = false, side_i = false;
            if (live) {
                side_i = t >= nj;
                m = side_i ? r0 + (t - nj) : jlist[j0 + t];
                sep = rflag[m] & RF_SEP;
                dirty = from != kNoMark && m > from;
            }
            const uint32_t hm = uint32_t(__ballot(sep) >> (32 * half));
            const uint32_t rank = c + uint32_t(__popc(hm & ((1u << sub) - 1u)));
            if (sep) mlist[mo + rank] = m | (side_i ? 0x80000000u : 0u);
            c += uint32_t(__popc(hm));
            const uint32_t pos = wave_reserve(dirty ? 1u : 0u, rx + RX_CNT0 + (cur ^ 1u), cap, rx + RX_OVER);
            if (dirty && pos != kNoMark) dnext[pos] = m;
        }
        if (act && sub == 0) {
            mcnt[v] = c;
            vchg[v] = kNoMark;
        }
    }
}

// The fixed point's moves, per voxel in pair order, on positions and
// previous positions; wlast / wfirst as the window walker's next pass reads
// them (moved voxels), `changed` when any pair separated.
__global__ void k_rx_commit_voxels(uint32_t n, const uint32_t *__restrict__ mcnt, const uint32_t *__restrict__ moff,
                                   const uint32_t *__restrict__ mlist,
                                   const uint32_t *__restrict__ rflag, const float *__restrict__ rsi,
                                   const float *__restrict__ rsj, float *rec, float *prev, uint32_t *wfirst,
                                   uint32_t *wlast, uint32_t *ctl) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint32_t c = mcnt[v];
    wfirst[v] = kNoMark;
    wlast[v] = c ? 1u : 0u;
    if (!c) return;
    float p[3], q[3];
    for (int a = 0; a < 3; ++a) {
        p[a] = rec[size_t(v) * kRec + a];
        q[a] = prev[size_t(v) * 3 + a];
    }
    for (uint32_t t = 0; t < c; ++t) {
        const uint32_t e = mlist[moff[v] + t];
        const uint32_t m = e & 0x7FFFFFFFu;
        const bool side_i = e >> 31;
        const float s = side_i ? rsi[m] : rsj[m];
        add_move(p, rflag[m], s, side_i);
        add_move(q, rflag[m], s, side_i);
    }
    for (int a = 0; a < 3; ++a) {
        rec[size_t(v) * kRec + a] = p[a];
        prev[size_t(v) * 3 + a] = q[a];
    }
    ctl[C_CHANGED] = 1;
}
// First contacts (the earliest pass that finds one) and this pass's outcomes.
__global__ void k_rx_commit_pairs(uint32_t P, const uint32_t *__restrict__ rflag, uint8_t *code, uint8_t *sep_last) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= P) return;
    const uint32_t f = rflag[m];
    if (!code[m]) code[m] = uint8_t(f >> 8);
    sep_last[m] = (f & RF_SEP) ? 1 : 0;
}
__global__ void k_set_word(uint32_t *w, uint32_t value) { *w = value; }

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return uint32_t((n + per - 1) / per); }

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" int32_t vx_contact_set_method(int32_t method, uint32_t max_rounds) {
    clear_error();
    if (method != VX_CONTACT_AUTO && method != VX_CONTACT_WINDOWS)
        return fail(VX_ERR_INVALID_ARGUMENT, "contact method must be VX_CONTACT_AUTO or VX_CONTACT_WINDOWS");
    Context &ctx = context();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex);
    ctx.contact_method = method;
    ctx.contact_rounds = max_rounds ? max_rounds : kRxRounds;
    return VX_OK;
}

extern "C" int32_t vx_contact_project(uint32_t n, float *positions, float *previous, const float *rest_positions,
                                      const float *mass, const uint8_t *kinematic_flags, const uint8_t *target_axes,
                                      const uint32_t *bond_neighbor, vx_contact *contacts, uint64_t capacity,
                                      vx_contact_report *report) {
    clear_error();
    vx_contact_report rep;
    std::memset(&rep, 0, sizeof rep);
    if (report) *report = rep;
    if (n > 0 && (!positions || !previous || !rest_positions || !mass || !kinematic_flags || !bond_neighbor))
        return fail(VX_ERR_INVALID_ARGUMENT, "vx_contact_project: null input array");
    if (capacity > 0 && !contacts) return fail(VX_ERR_INVALID_ARGUMENT, "vx_contact_project: null contacts");
    if (n == 0) {
        rep.passes = 1;  // one empty pass, as the reference's loop
        if (report) *report = rep;
        return VX_OK;
    }
    int32_t st = ensure_init();
    if (st != VX_OK) return st;
    Context &ctx = context();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex);
    hipStream_t s = ctx.stream;
    const size_t N = n;

    float *d_pos, *d_prev, *d_start, *d_rest, *d_mass;
    float4 *d_rec;
    uint8_t *d_flags;
    uint32_t *d_bond, *d_adj, *d_err;
    int32_t *d_lohi;
    uint64_t *d_cnt, *d_ends;
    unsigned long long *d_cand;
#define VX_BUF(name, bytes, ptr)                                               \
    do {                                                                       \
        int32_t b_ = buffer(name, (bytes), reinterpret_cast<void **>(&(ptr))); \
        if (b_ != VX_OK) return b_;                                            \
    } while (0)
    VX_BUF("contact.pos", 12 * N, d_pos);
    VX_BUF("contact.prev", 12 * N, d_prev);
    VX_BUF("contact.start", 12 * N, d_start);
    VX_BUF("contact.rest", 12 * N, d_rest);
    VX_BUF("contact.mass", 4 * N, d_mass);
    VX_BUF("contact.flags", N, d_flags);
    VX_BUF("contact.bond", 12 * N, d_bond);
    VX_BUF("contact.adj", 24 * N, d_adj);
    VX_BUF("contact.lohi", 24 * N, d_lohi);
    VX_BUF("contact.count", 8 * N, d_cnt);
    VX_BUF("contact.ends", 8 * N, d_ends);
    VX_BUF("contact.rec", 32 * N, d_rec);
    VX_BUF("contact.small", 64, d_err);
    d_cand = reinterpret_cast<unsigned long long *>(d_err + 8);

    std::vector<uint8_t> flags(N);
    for (size_t v = 0; v < N; ++v)
        flags[v] = uint8_t((kinematic_flags[v] ? 1 : 0) | ((target_axes ? (target_axes[v] & 7) : 0) << 1));
    VX_HIP(hipMemcpyAsync(d_pos, positions, 12 * N, hipMemcpyHostToDevice, s));
    VX_HIP(hipMemcpyAsync(d_prev, previous, 12 * N, hipMemcpyHostToDevice, s));
    VX_HIP(hipMemcpyAsync(d_rest, rest_positions, 12 * N, hipMemcpyHostToDevice, s));
    VX_HIP(hipMemcpyAsync(d_mass, mass, 4 * N, hipMemcpyHostToDevice, s));
    VX_HIP(hipMemcpyAsync(d_flags, flags.data(), N, hipMemcpyHostToDevice, s));
    VX_HIP(hipMemcpyAsync(d_bond, bond_neighbor, 12 * N, hipMemcpyHostToDevice, s));

    hipEvent_t ev[4];
    for (hipEvent_t &e : ev) VX_HIP(hipEventCreate(&e));
    struct EventGuard {
        hipEvent_t *e;
        ~EventGuard() {
            for (int q = 0; q < 4; ++q) (void)hipEventDestroy(e[q]);
        }
    } guard{ev};
    VX_HIP(hipEventRecord(ev[0], s));
    // let starts = scene.previous.clone() (src/collision.rs:36)
    VX_HIP(hipMemcpyAsync(d_start, d_prev, 12 * N, hipMemcpyDeviceToDevice, s));
    VX_HIP(hipMemsetAsync(d_adj, 0xFF, 24 * N, s));
    VX_HIP(hipMemsetAsync(d_err, 0, 64, s));
    uint32_t *d_box = d_err + 2;  // bin bounding box: minima (words 2-4), maxima (5-7)
    uint32_t *d_boxpart;
    const uint32_t nblocks = blocks_for(N, 256);
    VX_BUF("contact.boxpart", 24 * size_t(nblocks), d_boxpart);
    k_adjacency<<<blocks_for(3 * N, 256), 256, 0, s>>>(n, d_bond, d_adj, d_err);
    k_bins<<<nblocks, 256, 0, s>>>(n, d_start, d_pos, d_lohi, d_cnt, d_err, d_boxpart);
    k_box_reduce<<<1, 256, 0, s>>>(nblocks, d_boxpart, d_box);
    VX_HIP(hipGetLastError());

    // scratch for rocprim: grown to the largest request
    auto tmp = [&](size_t bytes, void **p) -> int32_t { return buffer("contact.tmp", bytes ? bytes : 16, p); };
    size_t tb = 0;
    void *d_tmp = nullptr;
    VX_HIP(rocprim::inclusive_scan(nullptr, tb, d_cnt, d_ends, N, rocprim::plus<uint64_t>(), s));
    if ((st = tmp(tb, &d_tmp)) != VX_OK) return st;
    VX_HIP(rocprim::inclusive_scan(d_tmp, tb, d_cnt, d_ends, N, rocprim::plus<uint64_t>(), s));

    uint32_t *pinned = static_cast<uint32_t *>(ctx.pinned);
    uint64_t *pin64 = reinterpret_cast<uint64_t *>(pinned + 64);
    VX_HIP(hipMemcpyAsync(pinned, d_err, 32, hipMemcpyDeviceToHost, s));
    VX_HIP(hipMemcpyAsync(pin64, d_ends + (N - 1), 8, hipMemcpyDeviceToHost, s));
    VX_HIP(hipStreamSynchronize(s));
    if (pinned[0] & 1) return fail(VX_ERR_INVALID_ARGUMENT, "vx_contact_project: bond neighbor index out of range");
    if (pinned[0] & 2)
        return fail(VX_ERR_INVALID_ARGUMENT,
                    "vx_contact_project: a swept box has no finite bound on some axis or reaches 2^20 cells from the origin");
    const uint64_t entries = pin64[0];
    if ((pinned[0] & 4) || entries > (uint64_t(1) << 31))
        return fail(VX_ERR_TOO_LARGE, "vx_contact_project: %llu bin entries (swept boxes too large)",
                    (unsigned long long)entries);
    // the bins' bounding box, and the bits its codes and the voxel indices take
    BinFrame frame{0, 0, 0, 1, 1};
    uint64_t cells = 1;
    if (entries > 0) {
        const uint32_t *bx = pinned + 2;
        frame.x0 = int(bx[0]) - kBinBias;
        frame.y0 = int(bx[1]) - kBinBias;
        frame.z0 = int(bx[2]) - kBinBias;
        frame.X = uint64_t(bx[3] - bx[0]) + 1;
        frame.Y = uint64_t(bx[4] - bx[1]) + 1;
        cells = frame.X * frame.Y * (uint64_t(bx[5] - bx[2]) + 1);
    }
    auto bit_width = [](uint64_t x) {
        int b = 0;
        while (b < 64 && (x >> b) != 0) ++b;
        return b;
    };
    const int bin_bits = std::max(1, bit_width(cells - 1));
    const uint32_t nb = uint32_t(std::max(1, bit_width(uint64_t(n) - 1)));

    uint64_t *d_keys, *d_keys2, *d_ecount, *d_eends;
    uint32_t *d_vals, *d_vals2;
    VX_BUF("contact.keys", 8 * entries, d_keys);
    VX_BUF("contact.keys2", 8 * entries, d_keys2);
    VX_BUF("contact.vals", 4 * entries, d_vals);
    VX_BUF("contact.vals2", 4 * entries, d_vals2);
    VX_BUF("contact.ecount", 8 * entries, d_ecount);
    VX_BUF("contact.eends", 8 * entries, d_eends);
    k_emit_bins<<<blocks_for(N, 256), 256, 0, s>>>(n, d_lohi, d_cnt, d_ends, d_keys, d_vals, frame);
    VX_HIP(hipGetLastError());
    tb = 0;
    VX_HIP(rocprim::radix_sort_pairs(nullptr, tb, d_keys, d_keys2, d_vals, d_vals2, size_t(entries), 0, bin_bits, s));
    if ((st = tmp(tb, &d_tmp)) != VX_OK) return st;
    VX_HIP(rocprim::radix_sort_pairs(d_tmp, tb, d_keys, d_keys2, d_vals, d_vals2, size_t(entries), 0, bin_bits, s));
    const uint32_t eblocks = blocks_for(entries, 256);
    unsigned long long *d_candpart;
    uint64_t *d_kmask;
    VX_BUF("contact.candpart", 8 * size_t(eblocks), d_candpart);
    VX_BUF("contact.kmask", 8 * entries, d_kmask);
    k_bin_pairs<false><<<eblocks, 256, 0, s>>>(entries, d_keys2, d_vals2, d_lohi, d_flags, d_rest, d_adj, d_ecount,
                                               nullptr, nullptr, d_candpart, frame, nb, d_kmask);
    k_sum_reduce<<<1, 256, 0, s>>>(eblocks, d_candpart, d_cand);
    VX_HIP(hipGetLastError());
    tb = 0;
    VX_HIP(rocprim::inclusive_scan(nullptr, tb, d_ecount, d_eends, size_t(entries), rocprim::plus<uint64_t>(), s));
    if ((st = tmp(tb, &d_tmp)) != VX_OK) return st;
    VX_HIP(rocprim::inclusive_scan(d_tmp, tb, d_ecount, d_eends, size_t(entries), rocprim::plus<uint64_t>(), s));
    VX_HIP(hipMemcpyAsync(pin64, d_eends + (entries - 1), 8, hipMemcpyDeviceToHost, s));
    VX_HIP(hipMemcpyAsync(pin64 + 1, d_cand, 8, hipMemcpyDeviceToHost, s));
    VX_HIP(hipStreamSynchronize(s));
    const uint64_t npairs = pin64[0];
    rep.candidate_pairs = pin64[1];
    rep.pairs = npairs;
    if (npairs >= (uint64_t(1) << 32))
        return fail(VX_ERR_TOO_LARGE, "vx_contact_project: %llu contact pairs", (unsigned long long)npairs);
    const uint32_t P = uint32_t(npairs);

    uint64_t *d_pk, *d_pairs, *d_cflag, *d_cends;
    uint8_t *d_code;
    vx_contact *d_out;
    VX_BUF("contact.pk", 8 * size_t(P) + 8, d_pk);
    VX_BUF("contact.pairs", 8 * size_t(P) + 8, d_pairs);
    VX_BUF("contact.code", size_t(P) + 8, d_code);
    VX_BUF("contact.cflag", 8 * size_t(P) + 8, d_cflag);
    VX_BUF("contact.cends", 8 * size_t(P) + 8, d_cends);
    uint64_t ncontacts = 0;
    if (P > 0) {
        k_bin_pairs<true><<<eblocks, 256, 0, s>>>(entries, d_keys2, d_vals2, d_lohi, d_flags, d_rest, d_adj, d_ecount,
                                                  d_eends, d_pk, nullptr, frame, nb, d_kmask);
        VX_HIP(hipGetLastError());
        tb = 0;
        VX_HIP(rocprim::radix_sort_keys(nullptr, tb, d_pk, d_pairs, size_t(P), 0, int(2 * nb), s));
        if ((st = tmp(tb, &d_tmp)) != VX_OK) return st;
        VX_HIP(rocprim::radix_sort_keys(d_tmp, tb, d_pk, d_pairs, size_t(P), 0, int(2 * nb), s));
        k_unpack_pairs<<<blocks_for(P, 256), 256, 0, s>>>(P, d_pairs, nb);
        uint8_t *d_sep, *d_due, *d_code_s, *d_sep_s;
        uint32_t *d_clist, *d_wfirst, *d_wlast, *d_moved, *d_ctl;
        float *d_pos_s, *d_prev_s;
        VX_BUF("contact.sep", size_t(P) + 8, d_sep);
        VX_BUF("contact.due", size_t(P) + 8, d_due);
        VX_BUF("contact.code_s", size_t(P) + 8, d_code_s);
        VX_BUF("contact.sep_s", size_t(P) + 8, d_sep_s);
        VX_BUF("contact.clist", 4 * size_t(P) + 8, d_clist);
        VX_BUF("contact.wfirst", 4 * N, d_wfirst);
        VX_BUF("contact.wlast", 4 * N, d_wlast);
        VX_BUF("contact.moved", 4 * N, d_moved);
        VX_BUF("contact.pos_s", 12 * N, d_pos_s);
        VX_BUF("contact.prev_s", 12 * N, d_prev_s);
        VX_BUF("contact.ctl", 4 * (C_WORDS + 1), d_ctl);
        uint32_t *d_ccount = d_ctl + C_WORDS;
        float *rec = reinterpret_cast<float *>(d_rec);
        VX_HIP(hipMemsetAsync(d_code, 0, P, s));
        VX_HIP(hipMemsetAsync(d_sep, 0, P, s));
        VX_HIP(hipMemsetAsync(d_wfirst, 0xFF, 4 * N, s));
        VX_HIP(hipMemsetAsync(d_wlast, 0, 4 * N, s));
        k_ctl_init<<<1, 1, 0, s>>>(d_ctl);
        k_pack<<<blocks_for(N, 256), 256, 0, s>>>(n, d_pos, d_start, d_mass, d_flags, d_rec);
        tb = 0;
        VX_HIP(rocprim::select(nullptr, tb, rocprim::counting_iterator<uint32_t>(0), d_due, d_clist, d_ccount,
                               size_t(P), s));
        if ((st = tmp(tb, &d_tmp)) != VX_OK) return st;
        VX_HIP(hipEventRecord(ev[2], s));
        const uint32_t vb = blocks_for(N, 256), pb = blocks_for(P, 256);
        // Passes by relaxation while they reach their fixed point within budget;
        // the window walker takes the first pass that does not, from its start.
        uint32_t relaxed = 0;  // passes done by relaxation
        bool finished = false;  // a relaxed pass moved nothing, or all passes are done
        if (ctx.contact_method == VX_CONTACT_AUTO && P < kRxMaxPairs) {
            uint32_t *d_rflag, *d_dl0, *d_dl1, *d_mcnt, *d_mlist, *d_moff, *d_mdeg, *d_vchg, *d_vstamp,
                *d_vlist, *d_jend, *d_joff, *d_jkey, *d_jkey2, *d_jlist, *d_pstart, *d_pend, *d_rx;
            float *d_rsi, *d_rsj;
            VX_BUF("contact.rx.rflag", 4 * size_t(P), d_rflag);
            VX_BUF("contact.rx.rsi", 4 * size_t(P), d_rsi);
            VX_BUF("contact.rx.rsj", 4 * size_t(P), d_rsj);
            VX_BUF("contact.rx.dl0", 4 * size_t(P), d_dl0);
            VX_BUF("contact.rx.dl1", 4 * size_t(P), d_dl1);
            VX_BUF("contact.rx.jlist", 4 * size_t(P), d_jlist);
            VX_BUF("contact.rx.jkey", 4 * size_t(P), d_jkey);
            VX_BUF("contact.rx.jkey2", 4 * size_t(P), d_jkey2);
            VX_BUF("contact.rx.mlist", 8 * size_t(P), d_mlist);
            VX_BUF("contact.rx.moff", 4 * N, d_moff);
            VX_BUF("contact.rx.mdeg", 4 * N, d_mdeg);
            VX_BUF("contact.rx.mcnt", 4 * N, d_mcnt);
            VX_BUF("contact.rx.vchg", 4 * N, d_vchg);
            VX_BUF("contact.rx.vstamp", 4 * N, d_vstamp);
            VX_BUF("contact.rx.vlist", 4 * N, d_vlist);
            VX_BUF("contact.rx.joff", 4 * N, d_joff);
            VX_BUF("contact.rx.jend", 4 * N, d_jend);
            VX_BUF("contact.rx.pstart", 4 * N, d_pstart);
            VX_BUF("contact.rx.pend", 4 * N, d_pend);
            VX_BUF("contact.rx.ctl", 64, d_rx);
            unsigned long long *d_evals = reinterpret_cast<unsigned long long *>(d_rx + 8);
            uint32_t *dl[2] = {d_dl0, d_dl1};
            k_rx_clear<<<blocks_for(std::max<uint64_t>(N, P), 256), 256, 0, s>>>(
                n, P, d_mcnt, d_vchg, d_vstamp, d_joff, d_jend, d_pstart, d_pend, d_rflag, d_rsi, d_rsj, d_rx);
            // the pairs of every voxel: (i, v) by a stable sort on v, (v, j) as ranges
            k_rx_jkeys<<<pb, 256, 0, s>>>(P, d_pairs, d_jkey);
            size_t sb = 0;
            VX_HIP(rocprim::radix_sort_pairs(nullptr, sb, d_jkey, d_jkey2, rocprim::counting_iterator<uint32_t>(0),
                                             d_jlist, size_t(P), 0, int(nb), s));
            void *d_stmp = nullptr;
            if ((st = buffer("contact.rx.tmp", sb ? sb : 16, &d_stmp)) != VX_OK) return st;
            VX_HIP(rocprim::radix_sort_pairs(d_stmp, sb, d_jkey, d_jkey2, rocprim::counting_iterator<uint32_t>(0),
                                             d_jlist, size_t(P), 0, int(nb), s));
            k_rx_bounds<<<pb, 256, 0, s>>>(P, d_pairs, d_jkey2, d_joff, d_jend, d_pstart, d_pend);
            // each voxel's mover list has room for all its pairs
            k_rx_degrees<<<vb, 256, 0, s>>>(n, d_joff, d_jend, d_pstart, d_pend, d_mdeg);
            size_t db = 0;
            VX_HIP(rocprim::exclusive_scan(nullptr, db, d_mdeg, d_moff, 0u, N, rocprim::plus<uint32_t>(), s));
            void *d_dtmp = nullptr;
            if ((st = buffer("contact.rx.tmp2", db ? db : 16, &d_dtmp)) != VX_OK) return st;
            VX_HIP(rocprim::exclusive_scan(d_dtmp, db, d_mdeg, d_moff, 0u, N, rocprim::plus<uint32_t>(), s));
            VX_HIP(hipGetLastError());
            uint32_t *pin_rx = pinned + 256;  // rx words (16: the evaluations, u64, at word 8), then ctl at 272
            uint32_t *pin_ctl = pinned + 272;
            uint32_t R = 0;                   // rounds enqueued so far; list R & 1 is the next round's
            uint32_t rx_rounds = 0;           // rounds that had dirty pairs
            uint64_t evals_before = 0;        // evaluations of the passes before this one
            // the first round of the first pass evaluates every pair; later rounds
            // are sized by the last dirty count the host has seen
            const uint32_t lb_all = std::min<uint32_t>(kRxBlocks, pb);
            auto grid_for = [&](uint64_t count) {
                return std::max<uint32_t>(64, std::min<uint32_t>(lb_all, blocks_for(count, 256)));
            };
            uint64_t seen = P;       // dirty pairs of the next round, as last seen
            bool pass_start = true;  // the next round is a pass's first
            uint64_t evals = 0;      // evaluations so far, as last seen
            auto enqueue_init = [&](uint32_t pass) {
                const uint32_t cur = R & 1u;
                if (pass == 0) {
                    k_rx_init_first<<<pb, 256, 0, s>>>(P, dl[cur], d_rx, cur);
                } else {
                    k_rx_init_words<<<1, 1, 0, s>>>(d_rx, cur);
                    k_rx_init_pass<<<blocks_for(N, 256), 256, 0, s>>>(n, d_mcnt, d_joff, d_jend, d_jlist, d_pstart,
                                                                    d_pend, d_rflag,
                                                      d_rsi, d_rsj, dl[cur], d_rx, cur, P);
                }
            };
            auto enqueue_commit = [&]() {
                k_set_word<<<1, 1, 0, s>>>(d_ctl + C_CHANGED, 0);
                k_rx_commit_voxels<<<vb, 256, 0, s>>>(n, d_mcnt, d_moff, d_mlist, d_rflag, d_rsi, d_rsj, rec, d_prev,
                                                      d_wfirst, d_wlast, d_ctl);
                k_rx_commit_pairs<<<pb, 256, 0, s>>>(P, d_rflag, d_code, d_sep);
            };
            // One host check per chunk of rounds. A pass that converges is
            // committed, and the next pass's start and first chunk are enqueued
            // before the next check, which then also reads whether the committed
            // pass changed anything: if not, the projection ended with it and the
            // pass just enqueued found no dirty pair (no separating pair, no
            // movers) and changed nothing.
            enqueue_init(0);
            uint32_t pass = 0;          // the pass being relaxed
            bool check_changed = false; // this check also answers whether pass - 1 changed anything
            for (;;) {
                const uint32_t left = ctx.contact_rounds - std::min(ctx.contact_rounds, rx_rounds);
                bool converged = false, over = false;
                if (left > 0) {
                    const uint32_t chunk = std::min(kRxChunk, left);
                    for (uint32_t q = 0; q < chunk; ++q, ++R) {
                        const uint32_t c = R & 1u;
                        // a pass's first round takes the whole list it starts from;
                        // the later ones take what the last check saw, or less
                        const uint32_t g = pass_start ? lb_all : std::min<uint32_t>(kRxBlocksLater, grid_for(seen));
                        pass_start = false;
                        k_rx_eval<<<g, 256, 0, s>>>(c, R + 1, dl[c], d_rx, P, d_pairs, rec, d_mcnt, d_moff, d_mlist,
                                                    d_rflag, d_rsi, d_rsj, d_vchg, d_vstamp, d_vlist, n, d_evals);
                        k_rx_rebuild<<<g, 256, 0, s>>>(c, d_rx, P, d_vlist, d_joff, d_jend, d_jlist, d_pstart, d_pend,
                                                       d_rflag, d_mcnt, d_moff, d_mlist, d_vchg, dl[c ^ 1u]);
                    }
                    VX_HIP(hipGetLastError());
                    VX_HIP(hipMemcpyAsync(pin_rx, d_rx, 16 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                    if (check_changed)
                        VX_HIP(hipMemcpyAsync(pin_ctl, d_ctl, 4 * C_WORDS, hipMemcpyDeviceToHost, s));
                    VX_HIP(hipStreamSynchronize(s));
                    std::memcpy(&evals, pin_rx + 8, 8);
                    rx_rounds = pin_rx[RX_ROUNDS];
                    over = pin_rx[RX_OVER] != 0;
                    seen = pin_rx[RX_CNT0 + (R & 1u)];
                    converged = !over && seen == 0;
                    if (check_changed) {
                        check_changed = false;
                        if (!pin_ctl[C_CHANGED]) {  // src/collision.rs:105-107
                            finished = true;
                            break;
                        }
                    }
                }
                if (!converged) {
                    // more rounds, or the walker does this pass from its start
                    // (`changed` still holds the previous pass's value, which the
                    // walker's next pass reads)
                    if (left > 0 && !over && evals - evals_before <= kRxEvalBudget * P) continue;
                    rep.evaluated += evals - evals_before;
                    break;
                }
                rep.evaluated += evals - evals_before;
                evals_before = evals;
                enqueue_commit();
                ++relaxed;
                if (relaxed == uint32_t(kPasses)) {
                    finished = true;
                    break;
                }
                ++pass;
                enqueue_init(pass);
                pass_start = true;
                check_changed = true;
            }
            rep.relaxation_rounds = rx_rounds;
            VX_HIP(hipGetLastError());
            rep.relaxed_passes = relaxed;
            if (!finished && relaxed > 0) k_set_word<<<1, 1, 0, s>>>(d_ctl + C_PASSES, relaxed);
        }
        if (finished) {
            k_set_word<<<1, 1, 0, s>>>(d_ctl + C_PASSES, relaxed);
            k_set_word<<<1, 1, 0, s>>>(d_ctl + C_DONE, 1);
        }
        if (!finished && relaxed == 0) {
            // the first pass's pairs before its first separating one, at once
            k_precheck<<<pb, 256, 0, s>>>(P, d_pairs, rec, d_code_s, d_ctl);
            k_first_pairs<<<pb, 256, 0, s>>>(P, d_ctl, d_code_s, d_code);
            k_first_done<<<1, 1, 0, s>>>(P, d_ctl);
            // the rest of pass 1 window by window
            k_project<<<1, kGs, 0, s>>>(d_pairs, nullptr, nullptr, P, rec, d_prev, d_code, d_sep, d_wfirst, d_wlast,
                                        d_ctl, 1);
            VX_HIP(hipGetLastError());
        }
        // passes after the walker's or the relaxation's last one, over candidates
        // (k_pass_begin stops them once a pass moved nothing)
        for (int pass = std::max<int>(1, int(relaxed)); !finished && pass < kPasses; ++pass) {
            k_pass_begin<<<1, 1, 0, s>>>(d_ctl);
            k_pass_voxels<<<vb, 256, 0, s>>>(n, d_ctl, d_moved, d_wfirst, d_wlast, rec, d_pos_s, d_prev, d_prev_s);
            k_pass_pairs<<<pb, 256, 0, s>>>(P, d_ctl, d_pairs, d_moved, d_sep, d_due, d_code, d_code_s, d_sep_s);
            k_choose<<<1, 1, 0, s>>>(P, d_ctl);
            for (int a = 0; a < kAttempts; ++a) {
                VX_HIP(rocprim::select(d_tmp, tb, rocprim::counting_iterator<uint32_t>(0), d_due, d_clist,
                                       d_ccount, size_t(P), s));
                k_project<<<1, kGs, 0, s>>>(d_pairs, d_clist, d_ccount, P, rec, d_prev, d_code, d_sep, d_wfirst,
                                            d_wlast, d_ctl);
                k_verify<<<pb, 256, 0, s>>>(P, d_ctl, d_pairs, d_due, d_wfirst);
                k_settle<<<1, 1, 0, s>>>(P, d_ctl);
                k_restore_voxels<<<vb, 256, 0, s>>>(n, d_ctl, rec, d_pos_s, d_prev, d_prev_s, d_wfirst, d_wlast);
                k_restore_pairs<<<pb, 256, 0, s>>>(P, d_ctl, d_code, d_code_s, d_sep, d_sep_s);
                k_restore_done<<<1, 1, 0, s>>>(d_ctl);
            }
            // no candidate pass verified: the pass over every pair (skipped when one did)
            k_project<<<1, kGs, 0, s>>>(d_pairs, nullptr, nullptr, P, rec, d_prev, d_code, d_sep, d_wfirst,
                                        d_wlast, d_ctl);
            VX_HIP(hipGetLastError());
        }
        VX_HIP(hipEventRecord(ev[3], s));
        k_unpack<<<blocks_for(N, 256), 256, 0, s>>>(n, reinterpret_cast<const float *>(d_rec), d_pos);
        k_contact_flags<<<blocks_for(P, 256), 256, 0, s>>>(P, d_code, d_cflag);
        tb = 0;
        VX_HIP(rocprim::inclusive_scan(nullptr, tb, d_cflag, d_cends, size_t(P), rocprim::plus<uint64_t>(), s));
        if ((st = tmp(tb, &d_tmp)) != VX_OK) return st;
        VX_HIP(rocprim::inclusive_scan(d_tmp, tb, d_cflag, d_cends, size_t(P), rocprim::plus<uint64_t>(), s));
        VX_HIP(hipMemcpyAsync(pin64, d_cends + (P - 1), 8, hipMemcpyDeviceToHost, s));
        VX_HIP(hipMemcpyAsync(pinned, d_ctl, 4 * C_WORDS, hipMemcpyDeviceToHost, s));
        VX_HIP(hipEventRecord(ev[1], s));
        VX_HIP(hipStreamSynchronize(s));
        ncontacts = pin64[0];
        rep.passes = pinned[C_PASSES];
        rep.rounds = pinned[C_WINDOWS];
        rep.redone = pinned[C_REDONE];
        rep.evaluated += pinned[C_EVALS];
    } else {
        VX_HIP(hipEventRecord(ev[1], s));
        VX_HIP(hipStreamSynchronize(s));
        rep.passes = 1;
    }
    float ms = 0.0f;
    VX_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    rep.device_us = double(ms) * 1000.0;
    if (P > 0) {
        VX_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
        rep.projection_us = double(ms) * 1000.0;
    }
    rep.contacts = ncontacts;
    if (report) *report = rep;
    if (ncontacts > capacity)
        return fail(VX_ERR_TOO_LARGE, "vx_contact_project: %llu contacts exceed the capacity %llu",
                    (unsigned long long)ncontacts, (unsigned long long)capacity);
    if (ncontacts > 0) {
        VX_BUF("contact.out", sizeof(vx_contact) * ncontacts, d_out);
        k_contact_scatter<<<blocks_for(P, 256), 256, 0, s>>>(P, d_pairs, d_code, d_cends, d_out);
        VX_HIP(hipGetLastError());
        VX_HIP(hipMemcpyAsync(contacts, d_out, sizeof(vx_contact) * ncontacts, hipMemcpyDeviceToHost, s));
    }
    VX_HIP(hipMemcpyAsync(positions, d_pos, 12 * N, hipMemcpyDeviceToHost, s));
    VX_HIP(hipMemcpyAsync(previous, d_prev, 12 * N, hipMemcpyDeviceToHost, s));
    VX_HIP(hipStreamSynchronize(s));
#undef VX_BUF
    return VX_OK;
}
