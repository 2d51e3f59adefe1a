// Device-resident structured scenes on a tile arena: rotation fit, bond
// assembly and the mixed-precision PCG solve for one or many REV boxes.
//
// HBM layout. An arena holds `cap_tiles` tiles of kBlock (512) REV rows;
// every field is a structure-of-arrays plane of R = cap_tiles * kBlock
// entries (plane q of a Q-wide field at base + q*R). A scene owns a
// contiguous range of tiles and its REV v = x + nx*(y + ny*z) lives at arena
// row voff + v. Per REV: positions / predicted / previous / velocity / drift
// (3 x f64), rotation (4 x f64), material constants (f32), flags (u8).
// Operator: per +axis bond a, k (symmetric 3x3, 6 values) in f64 and f32 and
// the bond force (3 x f64), stored on the bond's -end REV; per REV the
// projected diagonal D and its inverse (6 values, f64 and f32), the rhs
// (3 x f64) and the free mask (u8). Solve: the FP64 solution (two ping-pong
// buffers plus a warm-start copy) and the inner vectors d, r, z, p (x2), Ap
// in f32 (and f64 for the plain FP64 mode).
//
// The SpMV of a row reads its own three +axis blocks and the -axis
// neighbours' blocks (A_ji = A_ij^T = A_ij for a symmetric k); the free mask
// is applied to the row output and every PCG vector is zero on fixed DOFs.
//
// Scheduling. Kernels are launched over the whole arena once per tick; a
// workgroup maps to one tile, looks up the owning scene and acts only when
// that scene's phase bit for the kernel is set. Per-scene reductions sum the
// scene's per-tile partials (written with plain stores by the previous kernel)
// inside every workgroup, in a fixed order, so scalars are reproducible and
// need no atomics or fences. Tiles are interleaved over the 8 XCDs in chunks
// of `chunk` tiles (physical workgroup b runs on XCD b % 8).
//
// The algorithm restates src/solver.rs:27-110 (frames, compliance,
// stiffness, thermal rest), :327-417 (rotation fit), :1325-1345
// (constraint), :1552-1636 (assembly and projection), :475-570 and 668-683
// (substep, corotation loop with early exit, velocities); the PCG follows
// src/solver.rs:1833-1911 with block-Jacobi, run as FP32 inner solves inside
// an FP64 true-residual refinement loop.
#include "vx_common.hpp"
#include "vx_mech.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace vx {
namespace {

constexpr uint32_t kInnerChunk = 16;  // default inner iterations per tick (graph replay)
constexpr uint32_t kFreeTile = 0xFFFFFFFFu;

// Phase bits of a scene for one tick.
enum : uint32_t {
    B_PREDICT = 1u << 0,
    B_FIT = 1u << 1,      // fit after predict
    B_ASSEMBLE = 1u << 2,
    B_RESID0 = 1u << 3,   // first residual of a solve (x0 = 0 or warm start)
    // B_INNER: one chunk of inner iterations, preceded by the refinement
    // residual (x += d first) when the scene's inner solve has ended and its
    // outer loop has not; that decision is taken on the device.
    B_INNER = 1u << 5,
    B_APPLY = 1u << 6,
    B_FITB = 1u << 7,     // fit after apply, records the angle
    B_FINISH = 1u << 8,
    B_WARM_USE = 1u << 9,
    B_WARM_SAVE = 1u << 10,
    B_F64 = 1u << 11,     // inner solve in FP64 (plain FP64 mode)
    B_FRACTURE = 1u << 12,  // fracture pass of the substep, before the velocities
    B_RB = 1u << 13,        // red-black reduced system (red-black block SGS) instead of block-Jacobi
    B_TOOL = 1u << 14,      // the scene's rigid tool advances after the prediction
    B_MG = 1u << 15,        // aggregation multigrid cycle as the preconditioner (one-scene arenas, FP32 inner)
};

// Partial kinds; kFrac..kFrac+5 are the fracture pass's dissipated energies
// (plastic heat, stored plastic, surface work, damage heat, released) and
// its count of broken bonds.
enum : int { kPap = 0, kRz = 1, kRr = 2, kRr64 = 3, kBb = 4, kAngle = 5, kFrac = 6, kFracN = 6, kSlots = 12 };
enum : uint32_t { kRun = 0, kConv = 1, kNonFin = 2, kCap = 3 };

// Box geometry of a scene and its tile range; one copy per slot and one per
// tile (so a workgroup finds everything about its tile in one load).
struct SlotGeo {
    uint32_t nx, ny, nz, n;
    uint32_t tile0, ntiles, voff, slot;
};
struct SlotParams {
    double h, gx, gy, gz;
    double tol, inner_tol;
    uint32_t max_inner, max_outer;
    uint32_t fracture, pad;  // fracture: 1 when the scene runs the fracture pass
    FractureCfg fcfg;
};
// Written by the device; each field by one workgroup of one kernel and read
// by later kernels (and by the host after a tick).
struct SlotState {
    double rho[2], rr0, rr_last;         // inner PCG (rho by iteration parity)
    double rr64, bb, rel, angle;         // outer residual, corotation angle
    uint32_t k_spmv, k_update, iters, done;
    uint32_t sweeps, outer_done, xpar, inner_total;
    uint32_t nonfinite;
    uint32_t seg;   // iteration index at the last residual replacement
    uint32_t pbuf;  // p buffer holding the latest search direction
    uint32_t pad;
    double red[3];  // pre-summed p.Ap, r.z, r.r of a large scene (scene_reduce)
    double frac[kFracN];  // the last fracture pass: energies and broken bonds
};

static_assert(sizeof(SlotState) % 8 == 0, "SlotState is copied in 8-byte words");

// Scenes of more tiles than this have their inner partials summed once per
// kernel by scene_reduce instead of by every workgroup of the consumer (whose
// cost grows with the square of the tile count).
constexpr uint32_t kBigTiles = 1024;

struct DevArena {
    uint32_t R, cap_tiles, chunk, max_slots;
    const SlotGeo *tiles;  // per tile; slot == kFreeTile for a free tile
    const SlotGeo *geo;    // per slot
    const SlotParams *prm;
    uint32_t *bits;             // per slot, this tick
    uint32_t *tile_bits;        // per tile: the owning slot's bits (0 if free)
    SlotState *st;
    double *partials;  // kSlots * cap_tiles
    // state
    double *pos, *pred, *prev, *vel, *drift, *rot, *xa, *xb, *xw, *targets;
    float *mass, *young, *young_perp, *shear, *solid, *fiber, *aniso, *dtemp, *a_par, *a_perp, *loads;
    uint8_t *kinematic, *bonds, *tmask, *mask;
    float *damage, *kscale, *offset, *plastic;  // 3R, 3R, 9R ([a][c]), 9R
    // fracture: strength rows (12R), bond history (18R, [a][field]) and the
    // rotations of the last assembly (4R), which fix each bond's frame
    float *strength, *hist;
    double *rot_asm;
    // operator; after a substep's passes `force` holds each bond's last force
    // and `lambda` its multiplier (bond_forces)
    double *k64, *force, *d64, *dinv64, *rhs;  // 18R, 9R, 6R, 6R, 3R
    double *lambda;                             // 9R ([a][c])
    float *k32, *d32, *dinv32;                  // 18R, 6R, 6R
    // inner vectors
    float *d, *r, *z, *p0, *p1, *ap;
    double *dd, *rd, *zd, *p0d, *p1d, *apd;
    // With a list: workgroup L of a launch takes tile tile_list[L] (tiles of
    // the scenes in the tick, padded with kNoTile), so a tick of a few scenes
    // does not launch the tiles of the others.
    const uint32_t *tile_list;
};
constexpr uint32_t kNoTile = 0xFFFFFFFFu;

// ---- tile context -----------------------------------------------------------
// Physical workgroup b runs on XCD b % 8; logical tiles are dealt to the XCDs
// in chunks of `chunk` consecutive tiles.
__device__ __forceinline__ uint32_t logical_tile(uint32_t b, uint32_t chunk) {
    const uint32_t r = b / kXcds, x = b % kXcds;
    return ((r / chunk) * kXcds + x) * chunk + (r % chunk);
}

struct Tile {
    uint32_t t, slot, bits;
    SlotGeo g;
    uint32_t local, v;  // scene-local row and arena row of this thread
    bool row;
};

// Copy each slot's bits of this tick to its tiles, so a workgroup reads its
// tile's descriptor and bits side by side.
// Phase bits of up to kArgSlots scenes passed by value, so that a tick needs
// no host-to-device copy before its first kernel; n = 0: the bits are
// already in A.bits.
constexpr uint32_t kArgSlots = 256;
struct BitsArg {
    uint32_t n;
    uint32_t bits[kArgSlots];
};

// Per-slot bits (from the argument) and per-tile bits of this tick.
__global__ __launch_bounds__(kBlock) void spread_bits(DevArena A, BitsArg arg) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < arg.n) A.bits[t] = arg.bits[t];
    if (t >= A.cap_tiles) return;
    const uint32_t slot = A.tiles[t].slot;
    A.tile_bits[t] = slot == kFreeTile ? 0u : (arg.n ? arg.bits[slot] : A.bits[slot]);
}

// The tile of this workgroup and its scene; false for a free tile. The
// scene's phase bits are loaded but not tested.
__device__ __forceinline__ bool tile_base(const DevArena &A, Tile &c) {
    const uint32_t L = logical_tile(blockIdx.x, A.chunk);
    c.t = A.tile_list ? A.tile_list[L] : L;
    if (c.t >= A.cap_tiles) return false;
    c.g = A.tiles[c.t];
    c.bits = A.tile_bits[c.t];
    c.slot = c.g.slot;
    if (c.slot == kFreeTile) return false;
    c.local = (c.t - c.g.tile0) * kBlock + threadIdx.x;
    c.row = c.local < c.g.n;
    c.v = c.g.voff + c.local;
    return true;
}
__device__ __forceinline__ bool tile_for(const DevArena &A, uint32_t need, Tile &c) {
    return tile_base(A, c) && (c.bits & need);
}

__device__ __forceinline__ void coords(const SlotGeo &g, uint32_t local, uint32_t (&c)[3]) {
    c[0] = local % g.nx;
    const uint32_t t = local / g.nx;
    c[1] = t % g.ny;
    c[2] = t / g.ny;
}
__device__ __forceinline__ uint32_t stride_of(const SlotGeo &g, int a) { return a == 0 ? 1u : (a == 1 ? g.nx : g.nx * g.ny); }
__device__ __forceinline__ uint32_t extent_of(const SlotGeo &g, int a) { return a == 0 ? g.nx : (a == 1 ? g.ny : g.nz); }

__device__ __forceinline__ d3 ld3(const double *p, size_t R, uint32_t v) { return {p[v], p[R + v], p[2 * R + v]}; }
__device__ __forceinline__ void st3(double *p, size_t R, uint32_t v, d3 a) {
    p[v] = a.x;
    p[R + v] = a.y;
    p[2 * R + v] = a.z;
}
__device__ __forceinline__ q4 ldq(const double *p, size_t R, uint32_t v) { return {p[v], p[R + v], p[2 * R + v], p[3 * R + v]}; }
__device__ __forceinline__ Rev ld_rev(const DevArena &A, uint32_t v) {
    const size_t R = A.R;
    return Rev{A.young[v], A.young_perp[v], A.shear[v], A.solid[v], A.aniso[v],
               mk(A.fiber[v], A.fiber[R + v], A.fiber[2 * R + v]), A.dtemp[v], A.a_par[v], A.a_perp[v]};
}

// Per-scene sum of N partial kinds over the scene's tiles, broadcast to every
// thread of the workgroup (fixed order). scene_load is the per-thread part;
// block_sum_all completes it. inner_load is scene_load for the inner PCG
// partials, which a large scene takes pre-summed from its state.
// t: the thread's place in a 512-thread workgroup (a 256-thread workgroup
// with two rows per thread loads for t and t + 256).
template <int N>
__device__ __forceinline__ void scene_load(double (&out)[N], const DevArena &A, const SlotGeo &g, const int (&kind)[N],
                                           uint32_t t = threadIdx.x) {
    // One load per thread covers scenes of up to kBlock tiles without a loop
    // (a loop would wait for each load before the row loads are issued).
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double *part = A.partials + size_t(kind[k]) * A.cap_tiles + g.tile0;
        out[k] = t < g.ntiles ? part[t] : 0.0;
    }
    if (g.ntiles > kBlock) {
        // Larger scenes (scene_reduce): up to sixteen loads in flight per
        // thread and kind, added in the same order as one at a time.
        constexpr uint32_t kAhead = 16;
        for (uint32_t b = t + kBlock; b < g.ntiles; b += kAhead * kBlock) {
            double x[N][kAhead];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const double *part = A.partials + size_t(kind[k]) * A.cap_tiles + g.tile0;
#pragma unroll
                for (uint32_t u = 0; u < kAhead; ++u) x[k][u] = b + u * kBlock < g.ntiles ? part[b + u * kBlock] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < N; ++k)
#pragma unroll
                for (uint32_t u = 0; u < kAhead; ++u)
                    if (b + u * kBlock < g.ntiles) out[k] += x[k][u];
        }
    }
}
template <int N>
__device__ __forceinline__ void inner_load(double (&out)[N], const DevArena &A, const SlotGeo &g, const int (&kind)[N],
                                           uint32_t t = threadIdx.x) {
    if (g.ntiles > kBigTiles) {
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = t == 0 ? A.st[g.slot].red[kind[k]] : 0.0;
        return;
    }
    scene_load<N>(out, A, g, kind, t);
}
template <int N>
__device__ __forceinline__ void scene_sum(double (&out)[N], const DevArena &A, const SlotGeo &g, const int (&kind)[N],
                                          double (*lds)[kWaves]) {
    scene_load<N>(out, A, g, kind);
    block_sum_all<N>(out, lds);
}
template <int N>
__device__ __forceinline__ void tile_store(double (&v)[N], const DevArena &A, uint32_t t, const int (&kind)[N],
                                           double (*lds)[kWaves]) {
    block_sum_all<N>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) A.partials[size_t(kind[k]) * A.cap_tiles + t] = v[k];
    }
}

// ---- physics kernels -------------------------------------------------------
// Predict one substep (src/solver.rs:494-523, 542).
__global__ __launch_bounds__(kBlock) void substep_predict(DevArena A) {
    Tile c;
    if (!tile_for(A, B_PREDICT, c) || !c.row) return;
    const size_t R = A.R;
    const uint32_t v = c.v;
    const SlotParams p = A.prm[c.slot];
    d3 x = ld3(A.pos, R, v);
    st3(A.prev, R, v, x);
    const uint8_t tm = A.tmask[v];
    if (!A.kinematic[v]) {
        d3 vel = ld3(A.vel, R, v) + mk(p.gx, p.gy, p.gz) * p.h;
        if (tm & 1) vel.x = 0.0;
        if (tm & 2) vel.y = 0.0;
        if (tm & 4) vel.z = 0.0;
        st3(A.vel, R, v, vel);
        st3(A.drift, R, v, vel);
        x = x + vel * p.h;
    } else {
        st3(A.drift, R, v, mk(0, 0, 0));
    }
    if (tm & 1) x.x = A.targets[v];
    if (tm & 2) x.y = A.targets[R + v];
    if (tm & 4) x.z = A.targets[2 * R + v];
    st3(A.pos, R, v, x);
    st3(A.pred, R, v, x);
}

// Rotation fit (src/solver.rs:327-417) in FP64; with `record` the per-tile
// largest angle change goes to the kAngle partials.
__global__ __launch_bounds__(kBlock) void fit_rotations(DevArena A, uint32_t need, int record) {
    Tile c;
    if (!tile_for(A, need, c)) return;
    __shared__ double lds[1][kWaves];
    const size_t R = A.R;
    double angle = 0.0;
    if (c.row && !A.kinematic[c.v]) {
        const uint32_t v = c.v;
        uint32_t cc[3];
        coords(c.g, c.local, cc);
        const q4 prev = ldq(A.rot, R, v);
        const m3 old = mat_from_quat(prev);
        const d3 pv = ld3(A.pos, R, v);
        m3 f;
        d3 *cols[3] = {&f.c0, &f.c1, &f.c2};
        for (int a = 0; a < 3; ++a) {
            d3 sum{0, 0, 0};
            double cnt = 0.0;
            const uint32_t s = stride_of(c.g, a);
            if (cc[a] > 0 && (A.bonds[v - s] >> a & 1)) {
                sum = sum + (pv - ld3(A.pos, R, v - s)) * (1.0 / kCell);
                cnt += 1.0;
            }
            if (A.bonds[v] >> a & 1) {
                sum = sum + (ld3(A.pos, R, v + s) - pv) * (1.0 / kCell);
                cnt += 1.0;
            }
            *cols[a] = cnt > 0.0 ? sum * (1.0 / cnt) : col(old, a);
        }
        if (det3(f) > 1e-6) {
            for (int it = 0; it < 5; ++it) {
                m3 it_t = transpose3(inverse3m(f));
                m3 nx{(f.c0 + it_t.c0) * 0.5, (f.c1 + it_t.c1) * 0.5, (f.c2 + it_t.c2) * 0.5};
                double delta = 0.0;
                for (int a = 0; a < 3; ++a) {
                    d3 dd = col(nx, a) - col(f, a);
                    delta = fmax(delta, fmax(fmax(fabs(dd.x), fabs(dd.y)), fabs(dd.z)));
                }
                f = nx;
                if (delta <= 1e-6) break;
            }
            d3 ex = normalize_or_zero(f.c0);
            d3 ey = normalize_or_zero(f.c1 - ex * dot3(ex, f.c1));
            q4 q = prev;
            if (dot3(ex, ex) > 0.5 && dot3(ey, ey) > 0.5) {
                q = qnormalize(quat_from_mat(m3{ex, ey, cross3(ex, ey)}));
                A.rot[v] = q.x;
                A.rot[R + v] = q.y;
                A.rot[2 * R + v] = q.z;
                A.rot[3 * R + v] = q.w;
            }
            angle = 2.0 * acos(fmin(fmax(fabs(qdot(prev, q)), 0.0), 1.0));
        }
    }
    if (!record) return;
    angle = wave_max(angle);
    if ((threadIdx.x & 63) == 0) lds[0][threadIdx.x >> 6] = angle;
    lds_barrier();
    if (threadIdx.x == 0) {
        double m = lds[0][0];
        for (int w = 1; w < kWaves; ++w) m = fmax(m, lds[0][w]);
        A.partials[size_t(kAngle) * A.cap_tiles + c.t] = m;
    }
}

// Per +axis bond: rotated stiffness k and force k (p_j - p_i - rest); zero
// for absent bonds and bonds whose both endpoints are fully fixed
// (src/solver.rs:1595-1619).
__global__ __launch_bounds__(kBlock) void assemble_bonds(DevArena A) {
    Tile c;
    if (!tile_for(A, B_ASSEMBLE, c) || !c.row) return;
    const size_t R = A.R;
    const uint32_t v = c.v;
    const q4 qi = ldq(A.rot, R, v);
    const d3 pi = ld3(A.pos, R, v);
    const uint8_t mi = A.mask[v];
    const Rev ri = ld_rev(A, v);
    // The fracture pass and the bond forces evaluate every bond in the frame
    // of the last assembly (src/solver.rs:627-665).
    A.rot_asm[v] = qi.x;
    A.rot_asm[R + v] = qi.y;
    A.rot_asm[2 * R + v] = qi.z;
    A.rot_asm[3 * R + v] = qi.w;
    for (int a = 0; a < 3; ++a) {
        double k[6] = {0, 0, 0, 0, 0, 0};
        d3 force{0, 0, 0};
        const uint32_t j = v + stride_of(c.g, a);
        if ((A.bonds[v] >> a & 1) && (mi | A.mask[j])) {
            const size_t b = size_t(a) * R + v;
            const size_t o = size_t(a) * 3 * R;
            d3 rest;
            bond_constraint(a, ri, qi, ld_rev(A, j), ldq(A.rot, R, j),
                            double(A.kscale[b]) * fmax(1.0 - double(A.damage[b]), 0.001),
                            mk(A.offset[o + v], A.offset[o + R + v], A.offset[o + 2 * R + v]),
                            mk(A.plastic[o + v], A.plastic[o + R + v], A.plastic[o + 2 * R + v]), k, rest);
            d3 e = ld3(A.pos, R, j) - pi - rest;
            force = {k[0] * e.x + k[1] * e.y + k[2] * e.z, k[1] * e.x + k[3] * e.y + k[4] * e.z,
                     k[2] * e.x + k[4] * e.y + k[5] * e.z};
        }
        for (int q = 0; q < 6; ++q) {
            A.k64[(size_t(a) * 6 + q) * R + v] = k[q];
            A.k32[(size_t(a) * 6 + q) * R + v] = float(k[q]);
        }
        A.force[(size_t(a) * 3 + 0) * R + v] = force.x;
        A.force[(size_t(a) * 3 + 1) * R + v] = force.y;
        A.force[(size_t(a) * 3 + 2) * R + v] = force.z;
    }
}

// Per REV: D = m/h^2 I + sum k, rhs = (pred - p) m/h^2 + load +- forces in
// constraint order (-axis bond, then +axis bond, per axis), projection
// D <- P D P + I - P, rhs <- P rhs (src/solver.rs:1576-1636).
__global__ __launch_bounds__(kBlock) void assemble_rows(DevArena A) {
    Tile c;
    if (!tile_for(A, B_ASSEMBLE, c) || !c.row) return;
    const size_t R = A.R;
    const uint32_t v = c.v;
    const double h = A.prm[c.slot].h;
    uint32_t cc[3];
    coords(c.g, c.local, cc);
    const double m = double(A.mass[v]) / (h * h);
    double D[6] = {m, 0, 0, m, 0, m};
    d3 rhs = (ld3(A.pred, R, v) - ld3(A.pos, R, v)) * m;
    if (!A.kinematic[v]) rhs = rhs + mk(A.loads[v], A.loads[R + v], A.loads[2 * R + v]);
    for (int a = 0; a < 3; ++a) {
        if (cc[a] > 0) {
            const uint32_t u = v - stride_of(c.g, a);
            for (int q = 0; q < 6; ++q) D[q] += A.k64[(size_t(a) * 6 + q) * R + u];
            rhs = rhs - ld3(A.force + size_t(a) * 3 * R, R, u);
        }
        for (int q = 0; q < 6; ++q) D[q] += A.k64[(size_t(a) * 6 + q) * R + v];
        rhs = rhs + ld3(A.force + size_t(a) * 3 * R, R, v);
    }
    const uint8_t fm = A.mask[v];
    const double f[3] = {double(fm & 1), double(fm >> 1 & 1), double(fm >> 2 & 1)};
    const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double Dp[6];
    for (int r = 0; r < 3; ++r)
        for (int q = r; q < 3; ++q) Dp[idx[r][q]] = f[r] * D[idx[r][q]] * f[q] + (r == q ? 1.0 - f[r] : 0.0);
    double Di[6];
    sym_inverse(Dp, Di);
    for (int q = 0; q < 6; ++q) {
        A.d64[q * R + v] = Dp[q];
        A.d32[q * R + v] = float(Dp[q]);
        A.dinv64[q * R + v] = Di[q];
        A.dinv32[q * R + v] = float(Di[q]);
    }
    st3(A.rhs, R, v, mk(rhs.x * f[0], rhs.y * f[1], rhs.z * f[2]));
}

// ---- PCG kernels -------------------------------------------------------------
template <class S> __device__ __forceinline__ S fmaS(S a, S b, S c);
template <> __device__ __forceinline__ float fmaS<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fmaS<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class S>
__device__ __forceinline__ void symv(const S *__restrict__ m, size_t R, uint32_t v, const S (&x)[3], S (&y)[3]) {
    const S a0 = m[v], a1 = m[R + v], a2 = m[2 * R + v], a3 = m[3 * R + v], a4 = m[4 * R + v], a5 = m[5 * R + v];
    y[0] = fmaS<S>(a2, x[2], fmaS<S>(a1, x[1], a0 * x[0]));
    y[1] = fmaS<S>(a4, x[2], fmaS<S>(a3, x[1], a1 * x[0]));
    y[2] = fmaS<S>(a5, x[2], fmaS<S>(a4, x[1], a2 * x[0]));
}
template <class S>
__device__ __forceinline__ void symv_acc(const S *__restrict__ m, size_t R, uint32_t v, const S (&x)[3], S (&y)[3]) {
    const S a0 = m[v], a1 = m[R + v], a2 = m[2 * R + v], a3 = m[3 * R + v], a4 = m[4 * R + v], a5 = m[5 * R + v];
    y[0] = fmaS<S>(a2, x[2], fmaS<S>(a1, x[1], fmaS<S>(a0, x[0], y[0])));
    y[1] = fmaS<S>(a4, x[2], fmaS<S>(a3, x[1], fmaS<S>(a1, x[0], y[1])));
    y[2] = fmaS<S>(a5, x[2], fmaS<S>(a4, x[1], fmaS<S>(a2, x[0], y[2])));
}

template <class S> struct Inner;
template <> struct Inner<float> {
    static constexpr uint32_t kBit = 0;
    static __device__ __forceinline__ const float *k(const DevArena &a) { return a.k32; }
    static __device__ __forceinline__ const float *d(const DevArena &a) { return a.d32; }
    static __device__ __forceinline__ const float *dinv(const DevArena &a) { return a.dinv32; }
    static __device__ __forceinline__ float *dv(const DevArena &a) { return a.d; }
    static __device__ __forceinline__ float *r(const DevArena &a) { return a.r; }
    static __device__ __forceinline__ float *z(const DevArena &a) { return a.z; }
    static __device__ __forceinline__ float *p(const DevArena &a, int i) { return i ? a.p1 : a.p0; }
    static __device__ __forceinline__ float *ap(const DevArena &a) { return a.ap; }
};
template <> struct Inner<double> {
    static constexpr uint32_t kBit = B_F64;
    static __device__ __forceinline__ const double *k(const DevArena &a) { return a.k64; }
    static __device__ __forceinline__ const double *d(const DevArena &a) { return a.d64; }
    static __device__ __forceinline__ const double *dinv(const DevArena &a) { return a.dinv64; }
    static __device__ __forceinline__ double *dv(const DevArena &a) { return a.dd; }
    static __device__ __forceinline__ double *r(const DevArena &a) { return a.rd; }
    static __device__ __forceinline__ double *z(const DevArena &a) { return a.zd; }
    static __device__ __forceinline__ double *p(const DevArena &a, int i) { return i ? a.p1d : a.p0d; }
    static __device__ __forceinline__ double *ap(const DevArena &a) { return a.apd; }
};

template <class S>
__device__ __forceinline__ bool inner_bits(uint32_t bits) {
    return (bits & B_INNER) && ((bits & B_F64) == Inner<S>::kBit);
}

// A one-scene arena's inner kernels: the scene's geometry and the order in
// which workgroups visit its tiles (its launch shape). Tiles are visited by
// groups of zc = 2^zs tile columns that run through all z: with q tiles per
// z-plane (rounded), logical index L = ((jb * nzt) + zr) * zc + j visits tile
// zr * q + jb * zc + j, so a tile's +-z neighbour tiles are visited zc
// indices before and after it (the same XCD) instead of a whole plane away,
// and stay in that XCD's L2. q = 0 keeps the plain order. The shape changes
// which workgroup runs which tile, never what a tile computes: results are
// the same bit for bit under every shape.
constexpr uint32_t kZColsDefault = 8;
struct OneScene {
    SlotGeo g;
    uint32_t q, nzt, chunk, grid;
    uint32_t rb;  // the scene's bits carry B_RB
    uint32_t zs;  // log2 of the z-column group
};

// EARLY (template argument of the inner kernels): test the scene's done flag
// before the row loads, so the launches left in a chunk after the inner
// solve has ended fetch nothing. Chosen for arenas holding scenes of more
// than kBigTiles tiles, where the saved traffic outweighs the wait for the
// flag; at smaller sizes that wait lengthens every latency-bound launch.

// zcols (a power of two, or 0 for the plain order) groups the tile columns;
// planes of fewer than 16 tiles keep the plain order (their z neighbours are
// already within a few tiles).
inline OneScene make_one(const SlotGeo &g, bool round_robin, uint32_t zcols) {
    OneScene o{};
    o.g = g;
    const uint64_t plane = uint64_t(g.nx) * g.ny;
    const uint32_t q = uint32_t(std::max<uint64_t>(1, (plane + kBlock / 2) / kBlock));
    uint64_t space = g.ntiles;
    if (zcols && q >= 16) {
        o.q = q;
        o.zs = uint32_t(__builtin_ctz(zcols));
        o.nzt = (g.ntiles + q - 1) / q;
        space = uint64_t((q + zcols - 1) / zcols) * zcols * o.nzt;
    }
    // Logical tiles go to the XCDs either in contiguous eighths or round
    // robin (logical tile L on XCD L % 8, its z neighbours L +- zc on the
    // same XCD). Which is faster depends on the size (the tuned table
    // chooses per kernel group; the default deal is measured at
    // 192x192x96: the SpMV fastest with contiguous eighths, the kernels
    // after it round robin; at 64^3, where the working set stays in the
    // Infinity Cache, mixing the two mappings costs up to 7 %).
    if (round_robin) {
        o.chunk = 1;
        o.grid = uint32_t((space + kXcds - 1) / kXcds * kXcds);
    } else {
        o.chunk = uint32_t(std::max<uint64_t>(1, (space + kXcds - 1) / kXcds));
        o.grid = o.chunk * kXcds;
    }
    return o;
}

__device__ __forceinline__ uint32_t one_tile(uint32_t L, const OneScene &o) {
    if (o.q == 0) return L;
    const uint32_t j = L & ((1u << o.zs) - 1u), r = L >> o.zs, zr = r % o.nzt, col = ((r / o.nzt) << o.zs) + j;
    return col < o.q ? zr * o.q + col : kFreeTile;
}

// The tile of an inner-iteration kernel. ONE: the arena holds a single scene
// whose geometry comes in the kernel arguments (no descriptor load before
// the row loads), and the chunk is only launched while that scene iterates.
template <class S, bool ONE>
__device__ __forceinline__ bool inner_tile(const DevArena &A, const OneScene &one, Tile &c) {
    if (!ONE) return tile_base(A, c) && inner_bits<S>(c.bits);
    c.t = one_tile(logical_tile(blockIdx.x, one.chunk), one);
    if (c.t >= one.g.ntiles) return false;
    c.g = one.g;
    c.slot = one.g.slot;
    c.bits = B_INNER | Inner<S>::kBit | (one.rb ? uint32_t(B_RB) : 0u);
    c.local = c.t * kBlock + threadIdx.x;
    c.row = c.local < c.g.n;
    c.v = c.g.voff + c.local;
    return true;
}

// The head of an inner iteration, from the summed r.z (rho) and r.r of the
// previous update: false when the scene's inner solve has ended, else beta.
// A segment (the iterations since the last FP64 residual replacement) ends
// when its recurrence residual has dropped by inner_tol, or when it already
// meets the outer tolerance; the FP64 residual then decides. The lead thread
// records the iteration in the scene's state (one thread of one kernel per
// iteration).
template <class S>
__device__ __forceinline__ bool inner_head(SlotState *st, const SlotParams &prm, const double (&sums)[2], uint32_t k,
                                           uint32_t seg, double rr0_s, double bb, const double (&rho_s)[2], bool lead,
                                           int odd, S &beta) {
    const double rho = sums[0], rr = sums[1];
    const double rr0 = k == seg ? rr : rr0_s;
    uint32_t done = kRun;
    if (!isfinite(rr))
        done = kNonFin;
    else if (!(rr > prm.inner_tol * prm.inner_tol * rr0) || !(rr > prm.tol * prm.tol * bb))
        done = kConv;
    else if (k >= prm.max_inner)
        done = kCap;
    if (lead) {
        st->rr_last = rr;
        if (k == seg) st->rr0 = rr;
    }
    if (done) {
        if (lead) {
            st->done = done;
            st->iters = k;
        }
        return false;
    }
    const double rho_prev = k == 0 ? 0.0 : rho_s[(k - 1) & 1];
    beta = (k == 0 || !(rho_prev > 1e-300)) ? S(0) : S(rho / rho_prev);
    if (lead) {
        st->rho[k & 1] = rho;
        st->k_update = k;
        st->pbuf = odd ? 0u : 1u;
    }
    return true;
}

// ---- red-black reduced system -------------------------------------------------
// With the REVs coloured by the parity of x + y + z (red even, black odd),
// every bond joins a red and a black row, so the red-red and black-black parts
// of A are the block diagonals D_rr and D_bb. Eliminating the red rows leaves
// S x_b = b_b - A_br D_rr^-1 b_r with S = D_bb - A_br D_rr^-1 A_rb. CG on S
// with block-Jacobi (D_bb) is CG with red-black block symmetric Gauss-Seidel
// on A, in Eisenstat's form (the preconditioned operator is diag(I, S) in the
// scaled variables), at half the iterations of block-Jacobi on A. With the
// off-diagonal blocks of A equal to -k: A_rb p = -sum k p_nb, so for
// w_r = D_rr^-1 (mask * sum_nb k p_nb) the product is S p = D_bb p - mask *
// sum_nb k w_nb, the stencil of the SpMV with w in place of p at the
// neighbours. One iteration is three kernels: pcg_rb_red (w on the red
// rows), the SpMV of the black rows (S p, p.Ap), the update of the black
// rows. The inner vectors hold the black rows at colour-split positions
// (rb_pos); w holds the red rows at theirs, in the Ap planes. Operator
// planes (bond blocks, D, D^-1, mask) keep the natural layout. Every residual
// replacement first recovers
// x_r = D_rr^-1 (b_r + mask * sum_nb k x_nb) in FP64 (refine_red), so the FP64
// residual is the full system's.
__device__ __forceinline__ bool red_row(const uint32_t (&cc)[3]) { return ((cc[0] + cc[1] + cc[2]) & 1u) == 0u; }
__device__ __forceinline__ uint32_t colour_of(const SlotGeo &g, uint32_t v) {
    uint32_t cc[3];
    coords(g, v, cc);
    return red_row(cc) ? 0u : 1u;
}

// Colour-split positions of a red-black scene's inner vectors. Natural rows
// 2m and 2m + 1 of a box always differ in colour (they share an x-line of
// even length, or a z-plane of even size, or, with nx and ny odd, the
// colour is the parity of v itself), so m = v >> 1 numbers the rows of each
// colour in order. Row v of colour c sits at position rb_pos(c, v >> 1):
// blocks of 64 red rows then 64 black rows, so a wave reads one colour's
// values contiguously and a position stays within 64 of its natural row
// (the tile order's locality carries over). Positions of a scene of n rows
// stay below 128 * ceil(ceil(n / 2) / 64) <= its tile range.
__device__ __forceinline__ uint32_t rb_pos(uint32_t colour, uint32_t m) { return ((m >> 6) << 7) | (colour << 6) | (m & 63u); }

// D and D^-1 of a mixed-precision red-black scene at colour-split positions,
// so each colour's kernel reads them contiguously. They live in the FP64
// inner planes dd (D) and rd (D^-1), which an FP32 inner solve does not use:
// the six floats of position s fill the three doubles of row s, so a copy
// stays inside its own row's slots (a batch's FP64 scenes keep theirs).
// Refreshed by refine_red at the first residual of every solve.
__device__ __forceinline__ void ld_sym_rb(const double *plane, size_t R, uint32_t s, float (&o)[6]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const uint64_t t = ld_row(reinterpret_cast<const uint64_t *>(plane) + size_t(j) * R, s * 8u);
        o[2 * j] = __uint_as_float(uint32_t(t));
        o[2 * j + 1] = __uint_as_float(uint32_t(t >> 32));
    }
}
__device__ __forceinline__ void st_sym_rb(double *plane, size_t R, uint32_t s, const float (&v)[6]) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
        st_row(reinterpret_cast<uint64_t *>(plane) + size_t(j) * R, s * 8u,
               uint64_t(__float_as_uint(v[2 * j])) | (uint64_t(__float_as_uint(v[2 * j + 1])) << 32));
}

// The row a thread of a red-black inner kernel serves: its tile-local index s
// is a colour-split position; colour and natural row follow. False when the
// position holds no row of the wanted colour (wave-uniform colour).
struct RbRow {
    uint32_t colour, v;  // v: natural scene-local row
    bool row;
};
__device__ __forceinline__ RbRow rb_row(const SlotGeo &g, uint32_t s) {
    RbRow r;
    r.colour = (s >> 6) & 1u;
    const uint32_t m = ((s >> 7) << 6) | (s & 63u);
    const uint32_t v0 = 2u * m;
    r.v = v0 < g.n && colour_of(g, v0) == r.colour ? v0 : v0 + 1u;
    r.row = r.v < g.n;
    return r;
}

// Offsets of a red-black inner kernel's row: natural byte offsets of the row
// and its -axis neighbours (bond blocks, D, mask), colour-split byte offsets
// of the six neighbours (vectors); absent neighbours point at the row itself.
struct RbOffsets {
    uint32_t ov, om[3];   // natural
    uint32_t nb[6];       // colour-split, neighbour 2a (-axis a) and 2a + 1 (+axis a)
    bool lo[3], hi[3];
};
template <class S>
__device__ __forceinline__ RbOffsets rb_offsets(const SlotGeo &g, uint32_t v, uint32_t self_pos) {
    RbOffsets o;
    uint32_t cc[3];
    coords(g, v, cc);
    const uint32_t other = red_row(cc) ? 1u : 0u;
    const uint32_t sz = uint32_t(sizeof(S));
    o.ov = (g.voff + v) * sz;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t st = stride_of(g, a);
        o.lo[a] = cc[a] > 0;
        o.hi[a] = cc[a] + 1 < extent_of(g, a);
        o.om[a] = o.lo[a] ? o.ov - st * sz : o.ov;
        o.nb[2 * a] = (g.voff + (o.lo[a] ? rb_pos(other, (v - st) >> 1) : self_pos)) * sz;
        o.nb[2 * a + 1] = (g.voff + (o.hi[a] ? rb_pos(other, (v + st) >> 1) : self_pos)) * sz;
    }
    return o;
}

// sum over the six bonds of k u_nb (u zero where the neighbour is absent).
template <class S>
__device__ __forceinline__ void bond_sum(const S (&Km)[3][6], const S (&Kp)[3][6], const S (&u)[6][3], S (&acc)[3]) {
    acc[0] = acc[1] = acc[2] = S(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const S *km = Km[a], *kp = Kp[a];
        const S *pm = u[2 * a], *pq = u[2 * a + 1];
        acc[0] = fmaS<S>(km[2], pm[2], fmaS<S>(km[1], pm[1], fmaS<S>(km[0], pm[0], acc[0])));
        acc[1] = fmaS<S>(km[4], pm[2], fmaS<S>(km[3], pm[1], fmaS<S>(km[1], pm[0], acc[1])));
        acc[2] = fmaS<S>(km[5], pm[2], fmaS<S>(km[4], pm[1], fmaS<S>(km[2], pm[0], acc[2])));
        acc[0] = fmaS<S>(kp[2], pq[2], fmaS<S>(kp[1], pq[1], fmaS<S>(kp[0], pq[0], acc[0])));
        acc[1] = fmaS<S>(kp[4], pq[2], fmaS<S>(kp[3], pq[1], fmaS<S>(kp[1], pq[0], acc[1])));
        acc[2] = fmaS<S>(kp[5], pq[2], fmaS<S>(kp[4], pq[1], fmaS<S>(kp[2], pq[0], acc[2])));
    }
}

// The six bond blocks of a row (natural layout): its own +axis blocks and the
// -axis neighbours' (whatever they hold where the neighbour is absent; the
// vector they multiply is zero there).
template <class S>
__device__ __forceinline__ void ld_blocks(const S *__restrict__ K, size_t R, const RbOffsets &o, S (&Km)[3][6], S (&Kp)[3][6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const S *Ka = K + size_t(a) * 6 * R;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Km[a][q] = ld_row(Ka + q * R, o.om[a]);
            Kp[a][q] = ld_row(Ka + q * R, o.ov);
        }
    }
}

// The inner iteration's first kernel on the red rows: the inner check and
// beta (the SpMV takes the same from the same partials next, and records
// them), p = z + beta p_old formed at the six black neighbours, and
// w = D_rr^-1 (mask * sum k p_nb) into the red rows of the Ap planes. Waves
// of black positions only take part in the partial sums.
template <class S, bool ONE, bool EARLY>
__global__ __launch_bounds__(kBlock) void pcg_rb_red(DevArena A, OneScene one, int odd) {
    Tile c;
    if (!inner_tile<S, ONE>(A, one, c) || !(c.bits & B_RB)) return;
    __shared__ double lds[2][kWaves];
    SlotState *st = A.st + c.slot;
    const uint32_t done_in = st->done;
    if constexpr (EARLY) {
        if (done_in) return;  // the scene's inner solve has ended: no row loads
    }
    const uint32_t k = st->k_spmv;
    const uint32_t seg = st->seg;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const double rr0_s = st->rr0, bb = st->bb;
    const SlotParams prm = A.prm[c.slot];
    double sums[2];
    const int kinds[2] = {kRz, kRr};
    inner_load<2>(sums, A, c.g, kinds);
    __builtin_amdgcn_sched_barrier(0);
    const RbRow rr = rb_row(c.g, c.local);
    const bool wave_on = rr.colour == 0u;  // uniform over the wave
    const size_t R = A.R;
    const S *__restrict__ z = Inner<S>::z(A);
    const S *__restrict__ p_old = Inner<S>::p(A, odd);
    S zz[6][3], pp[6][3], Dv[6], Kp[3][6], Km[3][6];
    uint8_t m = 0;
    RbOffsets o{};
    if (wave_on) {
        o = rb_offsets<S>(c.g, rr.row ? rr.v : 0u, c.local);
        m = A.mask[o.ov / uint32_t(sizeof(S))];
        if constexpr (std::is_same<S, float>::value) {
            ld_sym_rb(A.rd, R, c.v, Dv);  // D^-1 at the own colour-split position
        } else {
            const S *__restrict__ Di = Inner<S>::dinv(A);
#pragma unroll
            for (int q = 0; q < 6; ++q) Dv[q] = ld_row(Di + q * R, o.ov);
        }
#pragma unroll
        for (int t = 0; t < 6; ++t)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                zz[t][q] = ld_row(z + q * R, o.nb[t]);
                pp[t][q] = ld_row(p_old + q * R, o.nb[t]);
            }
        ld_blocks<S>(Inner<S>::k(A), R, o, Km, Kp);
    }
    __builtin_amdgcn_sched_barrier(0);
    block_sum_all<2>(sums, lds);
    if (done_in || !wave_on) return;
    S beta;
    if (!inner_head<S>(st, prm, sums, k, seg, rr0_s, bb, rho_s, false, odd, beta)) return;
    S u[6][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            u[2 * a][q] = o.lo[a] ? fmaS<S>(beta, pp[2 * a][q], zz[2 * a][q]) : S(0);
            u[2 * a + 1][q] = o.hi[a] ? fmaS<S>(beta, pp[2 * a + 1][q], zz[2 * a + 1][q]) : S(0);
        }
    S sv[3];
    bond_sum<S>(Km, Kp, u, sv);
    if (!(m & 1)) sv[0] = S(0);
    if (!(m & 2)) sv[1] = S(0);
    if (!(m & 4)) sv[2] = S(0);
    if (rr.row) {
        S *__restrict__ w = Inner<S>::ap(A);
        const uint32_t os = c.v * uint32_t(sizeof(S));  // own colour-split position
        st_row(w, os, fmaS<S>(Dv[2], sv[2], fmaS<S>(Dv[1], sv[1], Dv[0] * sv[0])));
        st_row(w + R, os, fmaS<S>(Dv[4], sv[2], fmaS<S>(Dv[3], sv[1], Dv[1] * sv[0])));
        st_row(w + 2 * R, os, fmaS<S>(Dv[5], sv[2], fmaS<S>(Dv[4], sv[1], Dv[2] * sv[0])));
    }
}

// The SpMV of the reduced system on the black rows: the inner check and
// beta (recorded by the lead thread), p = z + beta p_old on the row,
// y = D p - mask * sum k w_nb with w from the red neighbours' Ap positions;
// stores p and y = S p at the row's position, partials of p.Ap. Waves of red
// positions only take part in the partial sums.
template <class S, bool EARLY>
__device__ __forceinline__ void spmv_black(const DevArena &A, const Tile &c, int odd, double (*lds)[kWaves]) {
    SlotState *st = A.st + c.slot;
    const uint32_t done_in = st->done;
    if constexpr (EARLY) {
        if (done_in) return;  // the scene's inner solve has ended: no row loads
    }
    const uint32_t k = st->k_spmv;
    const uint32_t seg = st->seg;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const double rr0_s = st->rr0, bb = st->bb;
    const SlotParams prm = A.prm[c.slot];
    double sums[2];
    const int kinds[2] = {kRz, kRr};
    inner_load<2>(sums, A, c.g, kinds);
    __builtin_amdgcn_sched_barrier(0);
    const RbRow rr = rb_row(c.g, c.local);
    const bool wave_on = rr.colour == 1u;
    const size_t R = A.R;
    const uint32_t os = (rr.row ? c.v : c.g.voff) * uint32_t(sizeof(S));  // own colour-split position
    const S *w = Inner<S>::ap(A);
    S zz[3], pp[3], ww[6][3], Dv[6], Kp[3][6], Km[3][6];
    uint8_t m = 0;
    RbOffsets o{};
    if (wave_on) {
        o = rb_offsets<S>(c.g, rr.row ? rr.v : 0u, rr.row ? c.local : 0u);
        const S *__restrict__ z = Inner<S>::z(A);
        const S *__restrict__ p_old = Inner<S>::p(A, odd);
        m = A.mask[o.ov / uint32_t(sizeof(S))];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            zz[q] = ld_row(z + q * R, os);
            pp[q] = ld_row(p_old + q * R, os);
        }
        if constexpr (std::is_same<S, float>::value) {
            ld_sym_rb(A.dd, R, os / uint32_t(sizeof(S)), Dv);  // D at the own colour-split position
        } else {
            const S *__restrict__ Dp = Inner<S>::d(A);
#pragma unroll
            for (int q = 0; q < 6; ++q) Dv[q] = ld_row(Dp + q * R, o.ov);
        }
#pragma unroll
        for (int t = 0; t < 6; ++t)
#pragma unroll
            for (int q = 0; q < 3; ++q) ww[t][q] = ld_row(w + q * R, o.nb[t]);
        ld_blocks<S>(Inner<S>::k(A), R, o, Km, Kp);
    }
    __builtin_amdgcn_sched_barrier(0);
    block_sum_all<2>(sums, lds);
    if (done_in) return;
    S beta;
    if (!inner_head<S>(st, prm, sums, k, seg, rr0_s, bb, rho_s, c.t == c.g.tile0 && threadIdx.x == 0, odd, beta)) return;
    double part[1] = {0.0};
    if (wave_on) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (!o.lo[a]) ww[2 * a][q] = S(0);
                if (!o.hi[a]) ww[2 * a + 1][q] = S(0);
            }
        S p[3], sv[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) p[q] = fmaS<S>(beta, pp[q], zz[q]);
        bond_sum<S>(Km, Kp, ww, sv);
        S y0 = fmaS<S>(Dv[2], p[2], fmaS<S>(Dv[1], p[1], Dv[0] * p[0]));
        S y1 = fmaS<S>(Dv[4], p[2], fmaS<S>(Dv[3], p[1], Dv[1] * p[0]));
        S y2 = fmaS<S>(Dv[5], p[2], fmaS<S>(Dv[4], p[1], Dv[2] * p[0]));
        if (m & 1) y0 -= sv[0];
        if (m & 2) y1 -= sv[1];
        if (m & 4) y2 -= sv[2];
        if (rr.row) {
            S *__restrict__ p_new = Inner<S>::p(A, !odd);
            S *ap = Inner<S>::ap(A);
            st_row(p_new, os, p[0]);
            st_row(p_new + R, os, p[1]);
            st_row(p_new + 2 * R, os, p[2]);
            st_row(ap, os, y0);
            st_row(ap + R, os, y1);
            st_row(ap + 2 * R, os, y2);
            part[0] = double(p[0]) * double(y0);
            part[0] = __builtin_fma(double(p[1]), double(y1), part[0]);
            part[0] = __builtin_fma(double(p[2]), double(y2), part[0]);
        }
    }
    const int out[1] = {kPap};
    tile_store<1>(part, A, c.t, out, lds);
}

// Inner convergence check of the previous update (r.r <= tol^2 r0.r0 every
// iteration), beta, then y = A p on the 7-point 3x3-block operator with
// p = z + beta p_old formed on load for the row and its neighbours; partials
// of p.Ap. `odd` selects the p ping-pong buffers.
//
// Issue order: the partials of the previous kernel first, then every row
// load (zeros where a neighbour is absent), then the partial sums. The row
// loads do not depend on beta, so they are in flight while the workgroup
// reduces the partials (LDS-only barriers do not wait for them).
template <class S, bool ONE, bool RB, bool EARLY>
__global__ __launch_bounds__(kBlock) void pcg_bond_spmv(DevArena A, OneScene one, int odd) {
    Tile c;
    if (!inner_tile<S, ONE>(A, one, c)) return;
    __shared__ double lds[2][kWaves];
    // RB: the launch may hold red-black scenes (all of them when ONE).
    if constexpr (RB) {
        if (ONE || (c.bits & B_RB)) {
            spmv_black<S, EARLY>(A, c, odd, lds);
            return;
        }
    }
    SlotState *st = A.st + c.slot;
    const uint32_t done_in = st->done;
    if constexpr (EARLY) {
        if (done_in) return;  // the scene's inner solve has ended: no row loads
    }
    const uint32_t k = st->k_spmv;
    const uint32_t seg = st->seg;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const double rr0_s = st->rr0, bb = st->bb;
    const SlotParams prm = A.prm[c.slot];
    double sums[2];
    const int kinds[2] = {kRz, kRr};
    inner_load<2>(sums, A, c.g, kinds);
    __builtin_amdgcn_sched_barrier(0);
    const size_t R = A.R;
    const S *__restrict__ z = Inner<S>::z(A);
    const S *__restrict__ p_old = Inner<S>::p(A, odd);
    const S *__restrict__ K = Inner<S>::k(A);
    const S *__restrict__ Dp = Inner<S>::d(A);
    const uint32_t v = c.row ? c.v : c.g.voff;  // rows past the scene load row 0 and discard
    uint32_t cc[3];
    coords(c.g, c.row ? c.local : 0u, cc);
    // Byte offsets of the row and its neighbours. An absent neighbour loads
    // the row itself (in bounds, cached, so no load sits behind a branch):
    // its +axis block is zero and its -axis block is zeroed below, so it
    // adds nothing.
    constexpr uint32_t sz = sizeof(S);
    const uint32_t ov = v * sz;
    uint32_t om[3], op[3];
    bool lo[3];
    // Large arenas (EARLY, FP32): the tile's own z and p_old are staged in
    // LDS, and a neighbour inside the same tile (x but at the tile's ends, y
    // when an x-line is shorter than the tile) is read from there instead of
    // through the caches: fewer memory requests per row. Rows past the scene
    // stage row 0's values; no row of the scene has them as a neighbour.
    constexpr bool kStage = EARLY && std::is_same<S, float>::value;
    __shared__ S stage[kStage ? 6 : 1][kStage ? kBlock : 1];
    bool inm[3] = {false, false, false}, inp[3] = {false, false, false};
    if constexpr (kStage) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = stride_of(c.g, a);
            const bool hi = cc[a] + 1 < extent_of(c.g, a);
            lo[a] = cc[a] > 0;
            om[a] = lo[a] ? ov - s * sz : ov;
            op[a] = hi ? ov + s * sz : ov;
            inm[a] = lo[a] && threadIdx.x >= s;
            inp[a] = hi && threadIdx.x + s < kBlock;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = stride_of(c.g, a);
            lo[a] = cc[a] > 0;
            om[a] = lo[a] ? ov - s * sz : ov;
            op[a] = cc[a] + 1 < extent_of(c.g, a) ? ov + s * sz : ov;
        }
    }
    S zz[7][3], pp[7][3], Dv[6], Kp[3][6], Km[3][6];
    const uint8_t m = A.mask[v];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        zz[0][q] = ld_row(z + q * R, ov);
        pp[0][q] = ld_row(p_old + q * R, ov);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) Dv[q] = ld_row(Dp + q * R, ov);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const S *Ka = K + size_t(a) * 6 * R;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (!inm[a]) {
                zz[1 + 2 * a][q] = ld_row(z + q * R, om[a]);
                pp[1 + 2 * a][q] = ld_row(p_old + q * R, om[a]);
            }
            if (!inp[a]) {
                zz[2 + 2 * a][q] = ld_row(z + q * R, op[a]);
                pp[2 + 2 * a][q] = ld_row(p_old + q * R, op[a]);
            }
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Km[a][q] = ld_row(Ka + q * R, om[a]);
            Kp[a][q] = ld_row(Ka + q * R, ov);  // zero when the +axis bond is absent
        }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int q = 0; q < 6; ++q) Km[a][q] = lo[a] ? Km[a][q] : S(0);
    if constexpr (kStage) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            stage[q][threadIdx.x] = zz[0][q];
            stage[3 + q][threadIdx.x] = pp[0][q];
        }
    }
    block_sum_all<2>(sums, lds);  // its barriers also order the staging stores before the reads below
    if (done_in) return;
    S beta;
    if (!inner_head<S>(st, prm, sums, k, seg, rr0_s, bb, rho_s, c.t == c.g.tile0 && threadIdx.x == 0, odd, beta)) return;
    if constexpr (kStage) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = stride_of(c.g, a);
            if (inm[a])
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    zz[1 + 2 * a][q] = stage[q][threadIdx.x - s];
                    pp[1 + 2 * a][q] = stage[3 + q][threadIdx.x - s];
                }
            if (inp[a])
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    zz[2 + 2 * a][q] = stage[q][threadIdx.x + s];
                    pp[2 + 2 * a][q] = stage[3 + q][threadIdx.x + s];
                }
        }
    }
    double part[1] = {0.0};
    S p[7][3];
#pragma unroll
    for (int t = 0; t < 7; ++t)
#pragma unroll
        for (int q = 0; q < 3; ++q) p[t][q] = fmaS<S>(beta, pp[t][q], zz[t][q]);
    S y0 = fmaS<S>(Dv[2], p[0][2], fmaS<S>(Dv[1], p[0][1], Dv[0] * p[0][0]));
    S y1 = fmaS<S>(Dv[4], p[0][2], fmaS<S>(Dv[3], p[0][1], Dv[1] * p[0][0]));
    S y2 = fmaS<S>(Dv[5], p[0][2], fmaS<S>(Dv[4], p[0][1], Dv[2] * p[0][0]));
    S a0 = S(0), a1 = S(0), a2 = S(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const S *km = Km[a], *kp = Kp[a];
        const S *pm = p[1 + 2 * a], *pq = p[2 + 2 * a];
        a0 = fmaS<S>(km[2], pm[2], fmaS<S>(km[1], pm[1], fmaS<S>(km[0], pm[0], a0)));
        a1 = fmaS<S>(km[4], pm[2], fmaS<S>(km[3], pm[1], fmaS<S>(km[1], pm[0], a1)));
        a2 = fmaS<S>(km[5], pm[2], fmaS<S>(km[4], pm[1], fmaS<S>(km[2], pm[0], a2)));
        a0 = fmaS<S>(kp[2], pq[2], fmaS<S>(kp[1], pq[1], fmaS<S>(kp[0], pq[0], a0)));
        a1 = fmaS<S>(kp[4], pq[2], fmaS<S>(kp[3], pq[1], fmaS<S>(kp[1], pq[0], a1)));
        a2 = fmaS<S>(kp[5], pq[2], fmaS<S>(kp[4], pq[1], fmaS<S>(kp[2], pq[0], a2)));
    }
    if (m & 1) y0 -= a0;
    if (m & 2) y1 -= a1;
    if (m & 4) y2 -= a2;
    if (c.row) {
        S *__restrict__ p_new = Inner<S>::p(A, !odd);
        S *__restrict__ ap = Inner<S>::ap(A);
        st_row(p_new, ov, p[0][0]);
        st_row(p_new + R, ov, p[0][1]);
        st_row(p_new + 2 * R, ov, p[0][2]);
        st_row(ap, ov, y0);
        st_row(ap + R, ov, y1);
        st_row(ap + 2 * R, ov, y2);
        part[0] = double(p[0][0]) * double(y0);
        part[0] = __builtin_fma(double(p[0][1]), double(y1), part[0]);
        part[0] = __builtin_fma(double(p[0][2]), double(y2), part[0]);
    }
    const int out[1] = {kPap};
    tile_store<1>(part, A, c.t, out, lds);
}

// The bond SpMV of large arenas (EARLY launches without red-black scenes):
// the same arithmetic, in the same order, as pcg_bond_spmv, with the
// neighbour loads streamed axis by axis instead of all at once. The own row,
// the x axis and the y axis's vectors are issued before the partial sums; the
// y axis's blocks and the z axis's vectors once the x axis has been consumed,
// the z axis's blocks after the y axis, and a scene's pre-summed r.z and r.r
// are read directly instead of through a workgroup sum. Fewer than two axes'
// values are live, so the kernel fits 6 waves per SIMD (three 512-thread
// workgroups per CU) instead of 4. FP32 only (the FP64 instance would spill).
// Measured in-solve: 96x96x64 20.5 -> 18.4 us per launch; 192x192x96 111.7
// -> 110.4 us (that launch is not occupancy-bound).
template <class S>
__device__ __forceinline__ void ld_axis_vec(const S *__restrict__ z, const S *__restrict__ p_old, size_t R, uint32_t om,
                                            uint32_t op, bool inm, bool inp, S (&zm)[3], S (&pm)[3], S (&zq)[3],
                                            S (&pq)[3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (!inm) {
            zm[q] = ld_row(z + q * R, om);
            pm[q] = ld_row(p_old + q * R, om);
        }
        if (!inp) {
            zq[q] = ld_row(z + q * R, op);
            pq[q] = ld_row(p_old + q * R, op);
        }
    }
}
template <class S>
__device__ __forceinline__ void ld_axis_blk(const S *__restrict__ Ka, size_t R, uint32_t om, uint32_t ov, S (&km)[6],
                                            S (&kp)[6]) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        km[q] = ld_row(Ka + q * R, om);
        kp[q] = ld_row(Ka + q * R, ov);  // zero when the +axis bond is absent
    }
}
// One axis of the bond sum: the -axis neighbour (block km, direction formed
// from zm, pm) then the +axis one, accumulated into a as pcg_bond_spmv does.
template <class S>
__device__ __forceinline__ void axis_sum(S beta, bool lo, const S (&zm)[3], const S (&pm)[3], const S (&zq)[3],
                                         const S (&pq)[3], const S (&km_in)[6], const S (&kp)[6], S &a0, S &a1, S &a2) {
    S km[6], u[3], w[3];
#pragma unroll
    for (int q = 0; q < 6; ++q) km[q] = lo ? km_in[q] : S(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        u[q] = fmaS<S>(beta, pm[q], zm[q]);
        w[q] = fmaS<S>(beta, pq[q], zq[q]);
    }
    a0 = fmaS<S>(km[2], u[2], fmaS<S>(km[1], u[1], fmaS<S>(km[0], u[0], a0)));
    a1 = fmaS<S>(km[4], u[2], fmaS<S>(km[3], u[1], fmaS<S>(km[1], u[0], a1)));
    a2 = fmaS<S>(km[5], u[2], fmaS<S>(km[4], u[1], fmaS<S>(km[2], u[0], a2)));
    a0 = fmaS<S>(kp[2], w[2], fmaS<S>(kp[1], w[1], fmaS<S>(kp[0], w[0], a0)));
    a1 = fmaS<S>(kp[4], w[2], fmaS<S>(kp[3], w[1], fmaS<S>(kp[1], w[0], a1)));
    a2 = fmaS<S>(kp[5], w[2], fmaS<S>(kp[4], w[1], fmaS<S>(kp[2], w[0], a2)));
}
template <class S, bool ONE, bool NT>
__global__ __launch_bounds__(kBlock, 6) void pcg_bond_spmv_big(DevArena A, OneScene one, int odd) {
    Tile c;
    if (!inner_tile<S, ONE>(A, one, c)) return;
    __shared__ double lds[2][kWaves];
    SlotState *st = A.st + c.slot;
    const uint32_t done_in = st->done;
    if (done_in) return;  // the scene's inner solve has ended: no row loads
    const uint32_t k = st->k_spmv;
    const uint32_t seg = st->seg;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const double rr0_s = st->rr0, bb = st->bb;
    const SlotParams prm = A.prm[c.slot];
    // A scene of more than kBigTiles tiles has its r.z and r.r pre-summed in
    // its state (scene_reduce): every thread reads them (uniform loads) and
    // no workgroup sum is needed. Smaller scenes of the same arena re-sum
    // their partials.
    const bool presummed = ONE || c.g.ntiles > kBigTiles;
    double sums[2];
    const int kinds[2] = {kRz, kRr};
    if (presummed) {
        sums[0] = st->red[kRz];
        sums[1] = st->red[kRr];
    } else {
        scene_load<2>(sums, A, c.g, kinds);
    }
    __builtin_amdgcn_sched_barrier(0);
    const size_t R = A.R;
    const S *__restrict__ z = Inner<S>::z(A);
    const S *__restrict__ p_old = Inner<S>::p(A, odd);
    const S *__restrict__ K = Inner<S>::k(A);
    const S *__restrict__ Dp = Inner<S>::d(A);
    const uint32_t v = c.row ? c.v : c.g.voff;  // rows past the scene load row 0 and discard
    uint32_t cc[3];
    coords(c.g, c.row ? c.local : 0u, cc);
    constexpr uint32_t sz = sizeof(S);
    const uint32_t ov = v * sz;
    // In-tile neighbours come from the staged own values in LDS (FP32).
    constexpr bool kStage = std::is_same<S, float>::value;
    __shared__ S stage[kStage ? 6 : 1][kStage ? kBlock : 1];
    uint32_t om[3], op[3], str[3];
    bool lo[3], inm[3] = {false, false, false}, inp[3] = {false, false, false};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        str[a] = stride_of(c.g, a);
        const bool hi = cc[a] + 1 < extent_of(c.g, a);
        lo[a] = cc[a] > 0;
        om[a] = lo[a] ? ov - str[a] * sz : ov;
        op[a] = hi ? ov + str[a] * sz : ov;
        if constexpr (kStage) {
            inm[a] = lo[a] && threadIdx.x >= str[a];
            inp[a] = hi && threadIdx.x + str[a] < kBlock;
        }
    }
    S z0[3], p0[3], Dv[6];
    S zm0[3], pm0[3], zq0[3], pq0[3], km0[6], kp0[6];
    S zm1[3], pm1[3], zq1[3], pq1[3], km1[6], kp1[6];
    S zm2[3], pm2[3], zq2[3], pq2[3], km2[6], kp2[6];
    const uint8_t m = A.mask[v];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        z0[q] = ld_row(z + q * R, ov);
        p0[q] = ld_row(p_old + q * R, ov);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) Dv[q] = NT ? ld_row_nt(Dp + q * R, ov) : ld_row(Dp + q * R, ov);  // read once
    ld_axis_vec<S>(z, p_old, R, om[0], op[0], inm[0], inp[0], zm0, pm0, zq0, pq0);
    ld_axis_blk<S>(K, R, om[0], ov, km0, kp0);
    ld_axis_vec<S>(z, p_old, R, om[1], op[1], inm[1], inp[1], zm1, pm1, zq1, pq1);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (kStage) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            stage[q][threadIdx.x] = z0[q];
            stage[3 + q][threadIdx.x] = p0[q];
        }
    }
    if (presummed) {
        if constexpr (kStage) __syncthreads();  // the staging stores before the reads below
    } else {
        block_sum_all<2>(sums, lds);  // its barriers also order the staging stores before the reads below
    }
    S beta;
    if (!inner_head<S>(st, prm, sums, k, seg, rr0_s, bb, rho_s, c.t == c.g.tile0 && threadIdx.x == 0, odd, beta)) return;
    auto from_stage = [&](int a, S (&zm)[3], S (&pm)[3], S (&zq)[3], S (&pq)[3]) {
        if constexpr (kStage) {
            if (inm[a])
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    zm[q] = stage[q][threadIdx.x - str[a]];
                    pm[q] = stage[3 + q][threadIdx.x - str[a]];
                }
            if (inp[a])
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    zq[q] = stage[q][threadIdx.x + str[a]];
                    pq[q] = stage[3 + q][threadIdx.x + str[a]];
                }
        }
    };
    S p[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) p[q] = fmaS<S>(beta, p0[q], z0[q]);
    S y0 = fmaS<S>(Dv[2], p[2], fmaS<S>(Dv[1], p[1], Dv[0] * p[0]));
    S y1 = fmaS<S>(Dv[4], p[2], fmaS<S>(Dv[3], p[1], Dv[1] * p[0]));
    S y2 = fmaS<S>(Dv[5], p[2], fmaS<S>(Dv[4], p[1], Dv[2] * p[0]));
    S a0 = S(0), a1 = S(0), a2 = S(0);
    from_stage(0, zm0, pm0, zq0, pq0);
    axis_sum<S>(beta, lo[0], zm0, pm0, zq0, pq0, km0, kp0, a0, a1, a2);
    // The x axis is summed before the next loads are issued (its registers
    // are free again): an empty asm that takes the sums and orders memory.
    asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2) : : "memory");
    ld_axis_blk<S>(K + 6 * R, R, om[1], ov, km1, kp1);
    ld_axis_vec<S>(z, p_old, R, om[2], op[2], inm[2], inp[2], zm2, pm2, zq2, pq2);
    __builtin_amdgcn_sched_barrier(0);
    from_stage(1, zm1, pm1, zq1, pq1);
    axis_sum<S>(beta, lo[1], zm1, pm1, zq1, pq1, km1, kp1, a0, a1, a2);
    asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2) : : "memory");
    ld_axis_blk<S>(K + 12 * R, R, om[2], ov, km2, kp2);
    from_stage(2, zm2, pm2, zq2, pq2);
    axis_sum<S>(beta, lo[2], zm2, pm2, zq2, pq2, km2, kp2, a0, a1, a2);
    if (m & 1) y0 -= a0;
    if (m & 2) y1 -= a1;
    if (m & 4) y2 -= a2;
    double part[1] = {0.0};
    if (c.row) {
        S *__restrict__ p_new = Inner<S>::p(A, !odd);
        S *__restrict__ ap = Inner<S>::ap(A);
        auto st_out = [](S *plane, uint32_t off, S val) {
            if constexpr (NT) st_row_nt(plane, off, val);
            else st_row(plane, off, val);
        };
        st_out(p_new, ov, p[0]);
        st_out(p_new + R, ov, p[1]);
        st_out(p_new + 2 * R, ov, p[2]);
        st_out(ap, ov, y0);
        st_out(ap + R, ov, y1);
        st_out(ap + 2 * R, ov, y2);
        part[0] = double(p[0]) * double(y0);
        part[0] = __builtin_fma(double(p[1]), double(y1), part[0]);
        part[0] = __builtin_fma(double(p[2]), double(y2), part[0]);
    }
    const int out[1] = {kPap};
    tile_store<1>(part, A, c.t, out, lds);
}

// alpha = rho / p.Ap; d += alpha p (d = alpha p on the first iteration of a
// segment);
// r -= alpha Ap; z = Dinv r; partials of r.z and r.r. Row loads are issued
// before the p.Ap partial sum, as in the SpMV.
template <class S, bool ONE, bool RB, bool EARLY, bool NT>
__global__ __launch_bounds__(kBlock) void pcg_update(DevArena A, OneScene one, int odd) {
    Tile c;
    if (!inner_tile<S, ONE>(A, one, c)) return;
    // A red-black scene updates its black rows at their colour-split
    // positions (the thread's own index); D^-1 stays in the natural layout.
    // Waves of red positions only take part in the partial sums.
    bool act = c.row, load = true, split_d = false;
    uint32_t v = c.row ? c.v : c.g.voff, vd = v;
    if constexpr (RB) {
        if (ONE || (c.bits & B_RB)) {
            const RbRow rr = rb_row(c.g, c.local);
            load = rr.colour == 1u;
            act = load && rr.row;
            v = c.v;
            vd = c.g.voff + (rr.row ? rr.v : 0u);
            split_d = std::is_same<S, float>::value;  // D^-1 at the own colour-split position
        }
    }
    __shared__ double lds[2][kWaves];
    SlotState *st = A.st + c.slot;
    const uint32_t done_in = st->done;
    if constexpr (EARLY) {
        if (done_in) return;  // the scene's inner solve has ended: no row loads
    }
    const uint32_t k = st->k_update;
    const uint32_t seg = st->seg;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    double pap[1];
    const int kinds[1] = {kPap};
    inner_load<1>(pap, A, c.g, kinds);
    __builtin_amdgcn_sched_barrier(0);
    const size_t R = A.R;
    const S *__restrict__ p = Inner<S>::p(A, !odd);
    const S *__restrict__ ap = Inner<S>::ap(A);
    S *__restrict__ d = Inner<S>::dv(A);
    S *__restrict__ r = Inner<S>::r(A);
    S *__restrict__ z = Inner<S>::z(A);
    const S *__restrict__ dinv = Inner<S>::dinv(A);
    const uint32_t ov = v * uint32_t(sizeof(S)), ovd = vd * uint32_t(sizeof(S));
    // NT: every plane streams (each is read once and written for a later
    // kernel; nothing is re-read within the launch).
    auto ld_once = [](const S *plane, uint32_t off) {
        if constexpr (NT) return ld_row_nt(plane, off);
        else return ld_row(plane, off);
    };
    auto st_out = [](S *plane, uint32_t off, S val) {
        if constexpr (NT) st_row_nt(plane, off, val);
        else st_row(plane, off, val);
    };
    S pv[3] = {}, av[3] = {}, rv[3] = {}, dv[3] = {}, Di[6] = {};
    if (load) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pv[q] = ld_once(p + q * R, ov);
            av[q] = ld_once(ap + q * R, ov);
            rv[q] = ld_once(r + q * R, ov);
            dv[q] = ld_once(d + q * R, ov);
        }
        if constexpr (RB && std::is_same<S, float>::value) {
            if (split_d) {
                ld_sym_rb(A.rd, R, v, Di);
            } else {
#pragma unroll
                for (int q = 0; q < 6; ++q) Di[q] = ld_once(dinv + q * R, ovd);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) Di[q] = ld_once(dinv + q * R, ovd);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    block_sum_all<1>(pap, lds);
    if (done_in) return;
    const S alpha = pap[0] > 1e-300 ? S(rho_s[k & 1] / pap[0]) : S(0);
    const bool first = k == seg;  // d restarts after each replacement
    if (c.t == c.g.tile0 && threadIdx.x == 0) {
        st->k_spmv = k + 1;
        st->iters = k + 1;
    }
    double part[2] = {0.0, 0.0};
    if (act) {
        S ri[3], zi[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            st_out(d + q * R, ov, fmaS<S>(alpha, pv[q], first ? S(0) : dv[q]));
            ri[q] = fmaS<S>(-alpha, av[q], rv[q]);
            st_out(r + q * R, ov, ri[q]);
        }
        zi[0] = fmaS<S>(Di[2], ri[2], fmaS<S>(Di[1], ri[1], Di[0] * ri[0]));
        zi[1] = fmaS<S>(Di[4], ri[2], fmaS<S>(Di[3], ri[1], Di[1] * ri[0]));
        zi[2] = fmaS<S>(Di[5], ri[2], fmaS<S>(Di[4], ri[1], Di[2] * ri[0]));
#pragma unroll
        for (int q = 0; q < 3; ++q) st_out(z + q * R, ov, zi[q]);
        part[0] = __builtin_fma(double(ri[2]), double(zi[2]), __builtin_fma(double(ri[1]), double(zi[1]), double(ri[0]) * double(zi[0])));
        part[1] = __builtin_fma(double(ri[2]), double(ri[2]), __builtin_fma(double(ri[1]), double(ri[1]), double(ri[0]) * double(ri[0])));
    }
    const int out[2] = {kRz, kRr};
    tile_store<2>(part, A, c.t, out, lds);
}

// pcg_update with two rows per thread: a 256-thread workgroup per 512-row
// tile of a block-Jacobi scene (a launch shape the autotuner may choose).
// Thread t holds rows t and t + 256, so wave w holds the 64-row groups w and
// w + 4; each group's wave sum goes to the slot the 512-thread form gives
// that group and the slots are added in the same order, and the partials
// re-summed at the head are split the same way: every value has the bits of
// the 512-thread form.
constexpr uint32_t kHalf = kBlock / 2;
template <int N>
__device__ __forceinline__ void block_sum_pair(double (&v0)[N], const double (&v1)[N], double (*lds)[kWaves]) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    lds_barrier();  // a previous use of lds is finished
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double s0 = wave_sum(v0[k]), s1 = wave_sum(v1[k]);
        if (lane == 0) {
            lds[k][wave] = s0;
            lds[k][wave + kWaves / 2] = s1;
        }
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = lds[k][0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += lds[k][w];
        v0[k] = s;
    }
}
template <class S, bool ONE, bool EARLY>
__global__ __launch_bounds__(kHalf) void pcg_update_pair(DevArena A, OneScene one, int odd) {
    Tile c;
    if (!inner_tile<S, ONE>(A, one, c)) return;
    const uint32_t l1 = c.local + kHalf;
    const bool act[2] = {c.row, l1 < c.g.n};
    const uint32_t v[2] = {c.row ? c.v : c.g.voff, act[1] ? c.g.voff + l1 : c.g.voff};
    __shared__ double lds[2][kWaves];
    SlotState *st = A.st + c.slot;
    const uint32_t done_in = st->done;
    if constexpr (EARLY) {
        if (done_in) return;  // the scene's inner solve has ended: no row loads
    }
    const uint32_t k = st->k_update;
    const uint32_t seg = st->seg;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    double pap[1], pap1[1];
    const int kinds[1] = {kPap};
    inner_load<1>(pap, A, c.g, kinds, threadIdx.x);
    inner_load<1>(pap1, A, c.g, kinds, threadIdx.x + kHalf);
    __builtin_amdgcn_sched_barrier(0);
    const size_t R = A.R;
    const S *__restrict__ p = Inner<S>::p(A, !odd);
    const S *__restrict__ ap = Inner<S>::ap(A);
    S *__restrict__ d = Inner<S>::dv(A);
    S *__restrict__ r = Inner<S>::r(A);
    S *__restrict__ z = Inner<S>::z(A);
    const S *__restrict__ dinv = Inner<S>::dinv(A);
    S pv[2][3], av[2][3], rv[2][3], dv[2][3], Di[2][6];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t ov = v[h] * uint32_t(sizeof(S));
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pv[h][q] = ld_row(p + q * R, ov);
            av[h][q] = ld_row(ap + q * R, ov);
            rv[h][q] = ld_row(r + q * R, ov);
            dv[h][q] = ld_row(d + q * R, ov);
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) Di[h][q] = ld_row(dinv + q * R, ov);
    }
    __builtin_amdgcn_sched_barrier(0);
    block_sum_pair<1>(pap, pap1, lds);
    if (done_in) return;
    const S alpha = pap[0] > 1e-300 ? S(rho_s[k & 1] / pap[0]) : S(0);
    const bool first = k == seg;  // d restarts after each replacement
    if (c.t == c.g.tile0 && threadIdx.x == 0) {
        st->k_spmv = k + 1;
        st->iters = k + 1;
    }
    double part[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (!act[h]) continue;
        const uint32_t ov = v[h] * uint32_t(sizeof(S));
        S ri[3], zi[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            st_row(d + q * R, ov, fmaS<S>(alpha, pv[h][q], first ? S(0) : dv[h][q]));
            ri[q] = fmaS<S>(-alpha, av[h][q], rv[h][q]);
            st_row(r + q * R, ov, ri[q]);
        }
        zi[0] = fmaS<S>(Di[h][2], ri[2], fmaS<S>(Di[h][1], ri[1], Di[h][0] * ri[0]));
        zi[1] = fmaS<S>(Di[h][4], ri[2], fmaS<S>(Di[h][3], ri[1], Di[h][1] * ri[0]));
        zi[2] = fmaS<S>(Di[h][5], ri[2], fmaS<S>(Di[h][4], ri[1], Di[h][2] * ri[0]));
#pragma unroll
        for (int q = 0; q < 3; ++q) st_row(z + q * R, ov, zi[q]);
        part[h][0] = __builtin_fma(double(ri[2]), double(zi[2]), __builtin_fma(double(ri[1]), double(zi[1]), double(ri[0]) * double(zi[0])));
        part[h][1] = __builtin_fma(double(ri[2]), double(ri[2]), __builtin_fma(double(ri[1]), double(ri[1]), double(ri[0]) * double(ri[0])));
    }
    block_sum_pair<2>(part[0], part[1], lds);
    if (threadIdx.x == 0) {
        A.partials[size_t(kRz) * A.cap_tiles + c.t] = part[0][0];
        A.partials[size_t(kRr) * A.cap_tiles + c.t] = part[0][1];
    }
}

// Whether the residual kernels act on a scene: `forced` launches run the
// first residual of a solve (B_RESID0); the launch at the head of an inner
// chunk runs the refinement residual once the inner solve has ended and the
// outer loop has not.
__device__ __forceinline__ bool resid_runs(uint32_t bits, const SlotState *st, int forced) {
    if (forced) return bits & B_RESID0;
    return (bits & B_INNER) && st->done != kRun && !st->outer_done;
}

// Red-black reduced system, before refine_residual (same `forced`): the new
// FP64 solution on every row of the scene. Black rows: x_new = x_old + d (or
// zero, or the warm start, on the first residual of a solve). Red rows, all
// of whose neighbours are black: x_new = D_rr^-1 (b + mask * sum_nb k x_nb),
// which zeroes the red rows of the residual b - A x.
template <class S>
__global__ __launch_bounds__(kBlock) void refine_red(DevArena A, int forced) {
    Tile c;
    if (!tile_base(A, c) || (c.bits & (B_F64 | B_RB)) != (Inner<S>::kBit | B_RB)) return;
    if (!resid_runs(c.bits, A.st + c.slot, forced) || !c.row) return;
    const size_t R = A.R;
    const SlotState *sst = A.st + c.slot;
    const uint32_t xpar = sst->xpar;
    const bool first = forced;
    const bool warm = first && (c.bits & B_WARM_USE);
    const bool use_d = !first && sst->iters > sst->seg;
    const double *__restrict__ x_old = warm ? A.xw : (xpar ? A.xb : A.xa);
    double *__restrict__ x_new = xpar ? A.xa : A.xb;
    const S *__restrict__ d = Inner<S>::dv(A);
    // x at black row u (arena row); d sits at u's colour-split position
    auto xval = [&](uint32_t u, double (&o)[3]) {
        const uint32_t du = c.g.voff + rb_pos(1u, (u - c.g.voff) >> 1);
#pragma unroll
        for (int q = 0; q < 3; ++q)
            o[q] = first ? (warm ? x_old[q * R + u] : 0.0) : x_old[q * R + u] + (use_d ? double(d[q * R + du]) : 0.0);
    };
    const uint32_t v = c.v;
    uint32_t cc[3];
    coords(c.g, c.local, cc);
    if constexpr (std::is_same<S, float>::value) {
        if (first) {
            // colour-split copies of D and D^-1 (assembled since the last solve)
            const uint32_t ps = c.g.voff + rb_pos(red_row(cc) ? 0u : 1u, c.local >> 1);
            float dv6[6], di6[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                dv6[q] = A.d32[q * R + v];
                di6[q] = A.dinv32[q * R + v];
            }
            st_sym_rb(A.dd, R, ps, dv6);
            st_sym_rb(A.rd, R, ps, di6);
        }
    }
    double xv[3];
    if (!red_row(cc)) {
        xval(v, xv);
    } else {
        double acc[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = stride_of(c.g, a);
            if (cc[a] > 0) {
                double xu[3];
                xval(v - s, xu);
                symv_acc<double>(A.k64 + size_t(a) * 6 * R, R, v - s, xu, acc);
            }
            if (cc[a] + 1 < extent_of(c.g, a)) {
                double xu[3];
                xval(v + s, xu);
                symv_acc<double>(A.k64 + size_t(a) * 6 * R, R, v, xu, acc);
            }
        }
        const uint8_t m = A.mask[v];
        double t[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) t[q] = A.rhs[q * R + v] + ((m >> q & 1) ? acc[q] : 0.0);
        symv<double>(A.dinv64, R, v, t, xv);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) x_new[q * R + v] = xv[q];
}

// FP64 residual replacement (mixed-precision PCG with reliable updates). The
// first residual of a solve starts from zero (or the warm start copy) and
// clears the search direction. A later one folds the segment's inner
// correction into the solution, x_new = x_old + d (own and neighbour rows),
// and keeps the search direction, moved to p buffer 0 where the next chunk's
// SpMV reads it. r64 = b - A64 x_new replaces the inner residual: r = r64 in
// the inner precision, z = Dinv r. Partials of ||r64||^2, ||b||^2, r.z, r.r.
// x_new goes to the other ping-pong buffer. A red-black scene's x_new is
// already complete (refine_red); its inner vectors hold the black rows (at
// their colour-split positions), so the inner partials cover the black rows
// while ||r64|| is the full system's.
template <class S>
__global__ __launch_bounds__(kBlock) void refine_residual(DevArena A, int forced) {
    Tile c;
    if (!tile_base(A, c) || ((c.bits & B_F64) != Inner<S>::kBit)) return;
    if (!resid_runs(c.bits, A.st + c.slot, forced)) return;
    __shared__ double lds[4][kWaves];
    const size_t R = A.R;
    const SlotState *sst = A.st + c.slot;
    const uint32_t xpar = sst->xpar;
    const bool first = forced;
    const bool warm = first && (c.bits & B_WARM_USE);
    // A segment that ended before its first update left d from the previous
    // segment, which x_old already holds.
    const bool use_d = !first && sst->iters > sst->seg;
    const uint32_t pbuf = sst->pbuf;
    const double *__restrict__ x_old = warm ? A.xw : (xpar ? A.xb : A.xa);
    double *__restrict__ x_new = xpar ? A.xa : A.xb;
    const S *__restrict__ d = Inner<S>::dv(A);
    const bool rb = c.bits & B_RB;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    if (c.row) {
        const uint32_t v = c.v;
        uint32_t cc[3];
        coords(c.g, c.local, cc);
        const bool red = rb && red_row(cc);
        auto xval = [&](uint32_t u, double (&o)[3]) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
                o[q] = rb ? x_new[q * R + u]
                          : (first ? (warm ? x_old[q * R + u] : 0.0) : x_old[q * R + u] + (use_d ? double(d[q * R + u]) : 0.0));
        };
        double xv[3];
        xval(v, xv);
        double yv[3];
        symv<double>(A.d64, R, v, xv, yv);
        double acc[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = stride_of(c.g, a);
            if (cc[a] > 0) {
                double xu[3];
                xval(v - s, xu);
                symv_acc<double>(A.k64 + size_t(a) * 6 * R, R, v - s, xu, acc);
            }
            if (cc[a] + 1 < extent_of(c.g, a)) {
                double xu[3];
                xval(v + s, xu);
                symv_acc<double>(A.k64 + size_t(a) * 6 * R, R, v, xu, acc);
            }
        }
        const uint8_t m = A.mask[v];
        const double f[3] = {double(m & 1), double(m >> 1 & 1), double(m >> 2 & 1)};
        S rs[3], zs[3];
        S *__restrict__ r = Inner<S>::r(A);
        S *__restrict__ z = Inner<S>::z(A);
        S *__restrict__ p0 = Inner<S>::p(A, 0);
        const S *__restrict__ p1 = Inner<S>::p(A, 1);
        // inner vectors of a red-black scene: black rows at their colour-split
        // positions, none for red rows
        const uint32_t pv = rb ? c.g.voff + rb_pos(1u, c.local >> 1) : v;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double b = A.rhs[q * R + v];
            const double r64 = f[q] * (b - (yv[q] - acc[q]));
            if (!rb) x_new[q * R + v] = xv[q];
            rs[q] = red ? S(0) : S(r64);
            if (!red) {
                r[q * R + pv] = rs[q];
                if (first)
                    p0[q * R + pv] = S(0);
                else if (pbuf)
                    p0[q * R + pv] = p1[q * R + pv];
            }
            part[0] = __builtin_fma(r64, r64, part[0]);
            part[1] = __builtin_fma(b, b, part[1]);
        }
        symv<S>(Inner<S>::dinv(A), R, v, rs, zs);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (!red) z[q * R + pv] = zs[q];
            part[2] = __builtin_fma(double(rs[q]), double(zs[q]), part[2]);
            part[3] = __builtin_fma(double(rs[q]), double(rs[q]), part[3]);
        }
    }
    const int out[4] = {kRr64, kBb, kRz, kRr};
    tile_store<4>(part, A, c.t, out, lds);
}

// One workgroup per scene after refine_residual (same `forced`): the outer
// relative residual, the decision (converged, or out of sweeps), the
// ping-pong flip and the start of the next inner segment. The iteration
// count k runs on across segments, so beta after a replacement is
// (r.z)_new / (r.z) of the last iteration.
template <class S>
__global__ __launch_bounds__(kBlock) void refine_decide(DevArena A, int forced) {
    const uint32_t slot = blockIdx.x;
    const uint32_t bits = A.bits[slot];
    SlotState *st = A.st + slot;
    if ((bits & B_F64) != Inner<S>::kBit || !resid_runs(bits, st, forced)) return;
    __shared__ double lds[2][kWaves];
    const SlotGeo g = A.geo[slot];
    if (g.n == 0) return;
    double s[2];
    const int kinds[2] = {kRr64, kBb};
    scene_sum<2>(s, A, g, kinds, lds);
    if (threadIdx.x != 0) return;
    const SlotParams p = A.prm[slot];
    const bool first = forced;
    const uint32_t sweeps = (first ? 0u : st->sweeps) + 1;
    const double rel = s[1] > 0.0 ? sqrt(s[0] / s[1]) : 0.0;
    const uint32_t k = first ? 0u : st->k_spmv;
    // Restart the search direction (beta = 0 at the next iteration) when the
    // true residual has drifted from the segment's recurrence residual,
    // ||r_true|| > 3 ||r_recurrence||: the old direction no longer fits it.
    if (!first && k > 0 && s[0] > 9.0 * st->rr_last) st->rho[(k - 1) & 1] = 0.0;
    st->inner_total = first ? 0u : st->iters;
    st->rr64 = s[0];
    st->bb = s[1];
    st->rel = rel;
    st->sweeps = sweeps;
    st->xpar ^= 1u;
    st->nonfinite = !isfinite(s[0]);
    const bool outer_done = !isfinite(s[0]) || rel <= p.tol || sweeps > p.max_outer;
    st->outer_done = outer_done;
    st->k_spmv = k;
    st->k_update = k;
    st->iters = k;
    st->seg = k;
    st->pbuf = 0;
    st->done = outer_done ? kConv : kRun;
}

// One workgroup per large scene: sums its partials of kinds [k0, k0 + nk)
// into red[] for the next inner kernel (forced: after the first residual of
// a solve; else inside an inner chunk).
template <class S>
__global__ __launch_bounds__(kBlock) void scene_reduce(DevArena A, int k0, int nk, int forced) {
    const uint32_t slot = blockIdx.x;
    const SlotGeo g = A.geo[slot];
    if (g.n == 0 || g.ntiles <= kBigTiles) return;
    const uint32_t bits = A.bits[slot];
    if ((bits & B_F64) != Inner<S>::kBit || !(bits & (forced ? uint32_t(B_RESID0) : uint32_t(B_INNER)))) return;
    __shared__ double lds[2][kWaves];
    SlotState *st = A.st + slot;
    if (nk == 2) {
        double s[2];
        const int kinds[2] = {k0, k0 + 1};
        scene_sum<2>(s, A, g, kinds, lds);
        if (threadIdx.x == 0) {
            st->red[k0] = s[0];
            st->red[k0 + 1] = s[1];
        }
    } else {
        double s[1];
        const int kinds[1] = {k0};
        scene_sum<1>(s, A, g, kinds, lds);
        if (threadIdx.x == 0) st->red[k0] = s[0];
    }
}

// positions += x on the free rows (src/solver.rs:1821-1823, 1921-1923); with
// B_WARM_SAVE the correction is also kept as the next solve's warm start.
__global__ __launch_bounds__(kBlock) void apply_correction(DevArena A) {
    Tile c;
    if (!tile_for(A, B_APPLY | B_WARM_SAVE, c) || !c.row) return;
    const size_t R = A.R;
    const uint32_t v = c.v;
    const double *x = A.st[c.slot].xpar ? A.xb : A.xa;
    const bool apply = (c.bits & B_APPLY) && A.mask[v];
    const bool save = c.bits & B_WARM_SAVE;
    for (int q = 0; q < 3; ++q) {
        const double xv = x[q * R + v];
        if (apply) A.pos[q * R + v] += xv;
        if (save) A.xw[q * R + v] = xv;
    }
}

// One workgroup per scene: largest rotation change of the fit after apply.
__global__ __launch_bounds__(kBlock) void reduce_angle(DevArena A) {
    const uint32_t slot = blockIdx.x;
    if (!(A.bits[slot] & B_FITB)) return;
    const SlotGeo g = A.geo[slot];
    if (g.n == 0) return;
    double m = 0.0;
    for (uint32_t b = threadIdx.x; b < g.ntiles; b += kBlock)
        m = fmax(m, A.partials[size_t(kAngle) * A.cap_tiles + g.tile0 + b]);
    __shared__ double lds[kWaves];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    lds_barrier();
    if (threadIdx.x == 0) {
        double best = lds[0];
        for (int w = 1; w < kWaves; ++w) best = fmax(best, lds[w]);
        A.st[slot].angle = best;
    }
}

// Fracture pass of a substep (src/solver.rs:627-665 with fracture::update,
// src/fracture.rs:210-292): for every intact +axis bond of a non-kinematic
// REV, after the corotation passes and before the velocities. The bond frame
// and the material stiffness (kscale included, damage not) are those of the
// last assembly; fibre alignment and the thermal rest length use the current
// rotations. A bond whose damage reaches one is removed from the lattice.
// Per tile: the dissipated energies and the count of broken bonds.
// After a substep's corotation passes, per constraint of the last assembly
// (every bond except those whose endpoints have no free axis and whose first
// REV is kinematic), at the solved positions:
//   lambda = -F^T k (p_j - p_i - rest) h^2 with the constraint's rotated
//            stiffness k and rest (src/solver.rs:1824-1827, 1921-1927);
//   last_force = F (K (u - plastic CELL)) (1 - damage) for a non-kinematic
//            first REV, u the elastic displacement in the bond frame
//            (src/solver.rs:627-640),
// both from the bond state before the fracture pass updates it, into the
// force and lambda planes (zero where there is no constraint).
__global__ __launch_bounds__(kBlock) void bond_forces(DevArena A) {
    Tile c;
    if (!tile_for(A, B_FINISH, c) || !c.row) return;
    const size_t R = A.R;
    const uint32_t v = c.v;
    const uint8_t bits = A.bonds[v];
    const bool kin = A.kinematic[v];
    const uint8_t mi = A.mask[v];
    const double h = A.prm[c.slot].h;
    const Rev ri = ld_rev(A, v);
    const q4 ai = ldq(A.rot_asm, R, v), qi = ldq(A.rot, R, v);
    const d3 pi = ld3(A.pos, R, v);
    for (int a = 0; a < 3; ++a) {
        d3 lam{0, 0, 0}, lf{0, 0, 0};
        const uint32_t j = v + stride_of(c.g, a);
        if ((bits >> a & 1) && !(kin && mi == 0 && A.mask[j] == 0)) {
            const size_t b = size_t(a) * R + v;
            const size_t o = size_t(a) * 3 * R;
            const Rev rj = ld_rev(A, j);
            const q4 aj = ldq(A.rot_asm, R, j), qj = ldq(A.rot, R, j);
            const d3 off = mk(A.offset[o + v], A.offset[o + R + v], A.offset[o + 2 * R + v]);
            const d3 pl = mk(A.plastic[o + v], A.plastic[o + R + v], A.plastic[o + 2 * R + v]);
            const double damage = A.damage[b];
            const m3 F = basis(a, slerp_half(ai, aj));
            const d3 pj = ld3(A.pos, R, j);
            double k[6];
            d3 rest;
            bond_constraint(a, ri, ai, rj, aj, double(A.kscale[b]) * fmax(1.0 - damage, 0.001), off, pl, k, rest);
            lam = mul_t(F, sym_mul(k, pj - pi - rest)) * (-h * h);
            if (!kin) {
                double K[6];
                local_stiffness(F, ri, ai, rj, aj, double(A.kscale[b]), K);
                const d3 u = mul_t(F, pj - pi) - thermal_rest_local(F, ri, qi, rj, qj) - off;
                lf = mul(F, sym_mul(K, u - pl * kCell)) * (1.0 - damage);
            }
        }
        for (int q = 0; q < 3; ++q) {
            A.force[(size_t(a) * 3 + q) * R + v] = q == 0 ? lf.x : (q == 1 ? lf.y : lf.z);
            A.lambda[(size_t(a) * 3 + q) * R + v] = q == 0 ? lam.x : (q == 1 ? lam.y : lam.z);
        }
    }
}

__global__ __launch_bounds__(kBlock) void fracture_bonds(DevArena A) {
    Tile c;
    if (!tile_for(A, B_FRACTURE, c)) return;
    __shared__ double lds[kFracN][kWaves];
    double acc[kFracN] = {0, 0, 0, 0, 0, 0};
    const size_t R = A.R;
    const uint32_t v = c.v;
    const uint8_t bits0 = c.row ? A.bonds[v] : uint8_t(0);
    if (c.row && bits0 && !A.kinematic[v]) {
        const SlotParams &p = A.prm[c.slot];
        const FractureCfg cfg = p.fcfg;
        const Rev ri = ld_rev(A, v);
        const q4 ai = ldq(A.rot_asm, R, v), qi = ldq(A.rot, R, v);
        const d3 pi = ld3(A.pos, R, v);
        float si[kStrength];
        for (int q = 0; q < kStrength; ++q) si[q] = A.strength[size_t(q) * R + v];
        const double solid_i = fmin(fmax(ri.solid, 0.0), 1.0);
        uint8_t bits = bits0;
        for (int a = 0; a < 3; ++a) {
            if (!(bits0 >> a & 1)) continue;
            const uint32_t j = v + stride_of(c.g, a);
            const size_t b = size_t(a) * R + v;
            const Rev rj = ld_rev(A, j);
            const q4 aj = ldq(A.rot_asm, R, j), qj = ldq(A.rot, R, j);
            const m3 F = basis(a, slerp_half(ai, aj));
            double K[6];
            local_stiffness(F, ri, ai, rj, aj, double(A.kscale[b]), K);
            float sj[kStrength];
            for (int q = 0; q < kStrength; ++q) sj[q] = A.strength[size_t(q) * R + j];
            const double solid = fmin(1.0, fmin(solid_i, fmin(fmax(rj.solid, 0.0), 1.0)));
            const Limits l = pair_limits(endpoint_limits(si, dot3(F.c0, qrotate(qi, ri.fiber)), ri.aniso),
                                         endpoint_limits(sj, dot3(F.c0, qrotate(qj, rj.fiber)), rj.aniso), solid * solid);
            const size_t o = size_t(a) * 3 * R;
            const d3 off = mk(A.offset[o + v], A.offset[o + R + v], A.offset[o + 2 * R + v]);
            const d3 u = mul_t(F, ld3(A.pos, R, j) - pi) - thermal_rest_local(F, ri, qi, rj, qj) - off;
            d3 plastic = mk(A.plastic[o + v], A.plastic[o + R + v], A.plastic[o + 2 * R + v]);
            double damage = A.damage[b];
            double h[kHistory];
            for (int f = 0; f < kHistory; ++f) h[f] = A.hist[(size_t(a) * kHistory + f) * R + v];
            Dissipation d;
            const bool broken = fracture_update(plastic, damage, h, u, K, l, p.h, cfg, d);
            A.plastic[o + v] = float(plastic.x);
            A.plastic[o + R + v] = float(plastic.y);
            A.plastic[o + 2 * R + v] = float(plastic.z);
            A.damage[b] = float(damage);
            for (int f = 0; f < kHistory; ++f) A.hist[(size_t(a) * kHistory + f) * R + v] = float(h[f]);
            for (int q = 0; q < 5; ++q) acc[q] += d.e[q];
            if (broken) {
                bits &= uint8_t(~(1u << a));
                acc[5] += 1.0;
            }
        }
        if (bits != bits0) A.bonds[v] = bits;
    }
    int kinds[kFracN];
    for (int q = 0; q < kFracN; ++q) kinds[q] = kFrac + q;
    tile_store<kFracN>(acc, A, c.t, kinds, lds);
}

// ---- rigid tool (one-scene arenas) -------------------------------------------
#include "vx_tool.hpp"

__device__ __forceinline__ t3 arena_pos(const DevArena &A, uint32_t v) {
    const size_t R = A.R;
    return t3{A.pos[v], A.pos[R + v], A.pos[2 * R + v]};
}

// The tool's candidate REVs, per tile: how many lie in its body box.
__global__ __launch_bounds__(kBlock) void tool_filter_count(DevArena A, ToolBufs B) {
    Tile c;
    if (!tile_for(A, B_TOOL, c)) return;
    __shared__ uint32_t wcnt[kWaves];
    const bool in = c.row && tool_in_box(*B.dev, arena_pos(A, c.v));
    const uint64_t m = __ballot(in);
    if ((threadIdx.x & 63u) == 0) wcnt[threadIdx.x >> 6] = uint32_t(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kWaves; ++w) t += wcnt[w];
        B.tile_count[c.t - c.g.tile0] = t;
    }
}

// Exclusive scan of the tiles' counts (one workgroup); the total is the
// candidate count.
__global__ __launch_bounds__(kToolBlock) void tool_filter_scan(ToolBufs B, uint32_t ntiles) {
    __shared__ uint32_t buf[kToolBlock];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += kToolBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < ntiles ? B.tile_count[i] : 0u;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t d = 1; d < kToolBlock; d <<= 1) {
            const uint32_t add = threadIdx.x >= d ? buf[threadIdx.x - d] : 0u;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < ntiles) B.tile_off[i] = carry + buf[threadIdx.x] - v;
        const uint32_t total = buf[kToolBlock - 1];
        __syncthreads();
        carry += total;
    }
    if (threadIdx.x == 0) {
        B.dev->candidates = carry;
        if (carry > kToolMaxCand) B.dev->status |= TS_CAND;
    }
}

// The candidates in index order: arena row and bin key.
__global__ __launch_bounds__(kBlock) void tool_filter_write(DevArena A, ToolBufs B) {
    Tile c;
    if (!tile_for(A, B_TOOL, c)) return;
    if (B.dev->candidates > kToolMaxCand) return;
    __shared__ uint32_t wcnt[kWaves];
    t3 p{0, 0, 0};
    bool in = false;
    if (c.row) {
        p = arena_pos(A, c.v);
        in = tool_in_box(*B.dev, p);
    }
    const uint64_t m = __ballot(in);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    if (lane == 0) wcnt[w] = uint32_t(__popcll(m));
    __syncthreads();
    if (!in) return;
    uint32_t at = B.tile_off[c.t - c.g.tile0] + uint32_t(__popcll(m & ((uint64_t(1) << lane) - 1)));
    for (uint32_t k = 0; k < w; ++k) at += wcnt[k];
    int32_t k3[3];
    uint64_t key = ~uint64_t(0);
    if (tool_key(p, k3))
        key = tool_pack(k3[0], k3[1], k3[2]);
    else
        atomicOr(&B.dev->status, uint32_t(TS_KEY));
    B.key[at] = key;
    B.idx[at] = c.v;
}

// One cell's bins in the reference's order (z, y, x), each bin's REVs in
// index order: fn(sorted candidate e, arena row v) for every one.
template <class F>
__device__ __forceinline__ void tool_cell_pairs(const ToolBufs &B, uint32_t C, t3 center, const t3 (&axes)[3], F &&fn) {
    int32_t lo[3], hi[3];
    if (!tool_bounds(center, axes, lo, hi)) return;
    for (int32_t z = lo[2]; z <= hi[2]; ++z)
        for (int32_t y = lo[1]; y <= hi[1]; ++y)
            for (int32_t x = lo[0]; x <= hi[0]; ++x) {
                const uint64_t key = tool_pack(x, y, z);
                uint32_t a = 0, b = C;  // first sorted candidate with this key
                while (a < b) {
                    const uint32_t mid = (a + b) >> 1;
                    if (B.key2[mid] < key)
                        a = mid + 1;
                    else
                        b = mid;
                }
                for (uint32_t e = a; e < C && B.key2[e] == key; ++e) fn(e, B.idx2[e]);
            }
}

// How far any cell of the tool is from its place at the advance's start (an
// upper bound: the centre's displacement plus the radius times the angle).
__device__ __forceinline__ double tool_moved(const ToolDev &T) {
#pragma clang fp contract(off)
    const tq q0 = tqld(T.rot0), q = tqld(T.rot);
    const double qd = fmin(fabs(((q0.x * q.x + q0.y * q.y) + q0.z * q.z) + q0.w * q.w), 1.0);
    return tlen(tld(T.pos) - tld(T.pos0)) + T.radius * 2.0 * acos(qd);
}

// Tool::advance for one substep (one workgroup; src/tool.rs:192-431); dt is
// the substep length, 0 (timed = 0) for the zero-time pass; g: the scene.
//
// The candidates cover every REV a bin of a cell can hold while the cells
// stay within `reach` of their starting place. Every pose change is checked
// against that before the next contact search. A tool that leaves it (the
// guided overlap correction moves a mounted tool by pen / (n . d) along its
// guide, large when a contact's normal is nearly across the guide) undoes the
// attempt -- the REV velocities it changed are restored, the tool state is
// the one at the advance's start -- and the workgroup gathers the candidates
// again with twice the reach, in index order as the filter kernels do, then
// runs the advance again: the result is the one a large enough box gives at
// once. After kToolAttempts the step fails (TS_MARGIN).
constexpr int kToolAttempts = 8;
__global__ __launch_bounds__(kToolBlock) void tool_advance(DevArena A, ToolBufs B, double dt, uint32_t timed, SlotGeo g) {
#pragma clang fp contract(off)
    __shared__ ToolDev T, T0;
    __shared__ double red_pen[kToolBlock / 64];
    __shared__ uint32_t red_cell[kToolBlock / 64];
    __shared__ uint32_t wcnt[kToolBlock / 64];
    __shared__ uint32_t done, lstatus, escaped, cand;
    if (threadIdx.x == 0) {
        T = *B.dev;
        T0 = T;
        lstatus = 0;
        escaped = 0;
        cand = T.candidates;
    }
    __syncthreads();
    if (T.status) return;
    const size_t R = A.R;
    const uint32_t nc = T.n_cells;
    uint32_t contacts = 0;  // thread 0
    for (int attempt = 0;; ++attempt) {
        const uint32_t C = cand;
        // Sort the candidates by (bin key, row): bitonic over the next power of two.
        uint32_t P = 1;
        while (P < C) P <<= 1;
        for (uint32_t i = threadIdx.x; i < P; i += kToolBlock) {
            B.key2[i] = i < C ? B.key[i] : ~uint64_t(0);
            B.idx2[i] = i < C ? B.idx[i] : ~0u;
        }
        __syncthreads();
        for (uint32_t k = 2; k <= P; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < P; i += kToolBlock) {
                    const uint32_t l = i ^ j;
                    if (l <= i) continue;
                    const uint64_t ka = B.key2[i], kb = B.key2[l];
                    const uint32_t ia = B.idx2[i], ib = B.idx2[l];
                    const bool greater = ka > kb || (ka == kb && ia > ib);
                    if (greater == ((i & k) == 0)) {
                        B.key2[i] = kb;
                        B.key2[l] = ka;
                        B.idx2[i] = ib;
                        B.idx2[l] = ia;
                    }
                }
                __syncthreads();
            }
        // The candidates' velocities at the start (an escaped attempt restores
        // them) and their contact velocities.
        for (uint32_t e = threadIdx.x; e < C; e += kToolBlock) {
            const uint32_t v = B.idx2[e];
            for (int q = 0; q < 3; ++q) B.vsave[3 * size_t(e) + q] = A.vel[q * R + v];
            if (timed)
                for (int q = 0; q < 3; ++q) B.cv[3 * size_t(e) + q] = 0.0;
        }
        if (threadIdx.x == 0) contacts = 0;
        __syncthreads();
        for (uint32_t step = 0; step < T.micro; ++step) {
            // Every thread is past its reads of `escaped` and `done` from the
            // previous microstep before thread 0 writes them again.
            __syncthreads();
            if (threadIdx.x == 0) {
                const double h = T.h;
                tst(T.vel, tld(T.vel) + tld(T.lin_acc) * h);
                tst(T.ang, tld(T.ang) + tool_angular(T, tld(T.torque)) * h);
                tst(T.pos, tld(T.pos) + tld(T.vel) * h);
                const double angle = tlen(tld(T.ang)) * h;
                if (angle > 1e-8) {
                    const t3 ax = tdiv(tld(T.ang), tlen(tld(T.ang)));
                    const double sn = sin(angle * 0.5), cs = cos(angle * 0.5);
                    const tq r{ax.x * sn, ax.y * sn, ax.z * sn, cs};
                    const tq q = tqld(T.rot);
                    // Hamilton product r * q, then normalise
                    tq m{r.w * q.x + r.x * q.w + r.y * q.z - r.z * q.y, r.w * q.y - r.x * q.z + r.y * q.w + r.z * q.x,
                         r.w * q.z + r.x * q.y - r.y * q.x + r.z * q.w, r.w * q.w - r.x * q.x - r.y * q.y - r.z * q.z};
                    const double len = sqrt(((m.x * m.x + m.y * m.y) + m.z * m.z) + m.w * m.w);
                    const double inv = 1.0 / len;
                    tqst(T.rot, tq{m.x * inv, m.y * inv, m.z * inv, m.w * inv});
                }
                if (tool_moved(T) > T.reach) escaped = 1;
            }
            __syncthreads();
            if (escaped) break;
            const tq rot = tqld(T.rot);
            const t3 axes[3] = {trot(rot, tv(1, 0, 0)), trot(rot, tv(0, 1, 0)), trot(rot, tv(0, 0, 1))};
            // Contacts of every cell, in order, in parallel.
            for (uint32_t c = threadIdx.x; c < nc; c += kToolBlock) {
                const t3 center = tld(T.pos) + trot(rot, tld(B.cells + 3 * c));
                uint32_t n = 0;
                tool_cell_pairs(B, C, center, axes, [&](uint32_t e, uint32_t v) {
                    t3 nrm;
                    double pen;
                    if (!tool_sat(center, axes, arena_pos(A, v), nrm, pen)) return;
                    if (n < kToolCellSlots) {
                        ToolContact &tc = B.con[size_t(c) * kToolCellSlots + n];
                        tc.slot = e;
                        tst(tc.n, nrm);
                        tc.pen = pen;
                    } else {
                        atomicOr(&lstatus, uint32_t(TS_SLOTS));
                    }
                    ++n;
                });
                B.ncon[c] = n < kToolCellSlots ? n : kToolCellSlots;
            }
            __syncthreads();
            // The impulses, sequentially in the reference's order.
            if (threadIdx.x == 0) {
                t3 correction{0, 0, 0};
                double max_pen = 0.0;
                for (uint32_t c = 0; c < nc; ++c) {
                    const uint32_t nn = B.ncon[c];
                    for (uint32_t q = 0; q < nn; ++q) {
                        const ToolContact tc = B.con[size_t(c) * kToolCellSlots + q];
                        const uint32_t e = tc.slot, v = B.idx2[e];
                        const t3 nrm = tld(tc.n);
                        contacts += 1;
                        const t3 p = arena_pos(A, v);
                        const t3 arm = p - tld(T.pos);
                        t3 vi{A.vel[v], A.vel[R + v], A.vel[2 * R + v]};
                        const t3 rel = tld(T.vel) + tcross(tld(T.ang), arm) - vi;
                        const double closing = tdot(rel, nrm);
                        const bool kin = A.kinematic[v] != 0;
                        const uint8_t tm = A.tmask[v];
                        const t3 fr{(!kin && !(tm & 1)) ? 1.0 : 0.0, (!kin && !(tm & 2)) ? 1.0 : 0.0,
                                    (!kin && !(tm & 4)) ? 1.0 : 0.0};
                        const double mi = double(A.mass[v]);
                        const double wi = !kin ? tdot(thmul(nrm, fr), nrm) / mi : 0.0;
                        const double denom =
                            wi + tdot(nrm, tool_linear(T, nrm)) + tdot(nrm, tcross(tool_angular(T, tcross(arm, nrm)), arm));
                        if (closing > 0.0) {
                            const double ni = closing / fmax(denom, 1e-9);
                            const t3 nv = nrm * ni;
                            const t3 dv = tdiv(thmul(nv, fr), mi);
                            vi = vi + dv;
                            if (timed) tst(B.cv + 3 * size_t(e), tld(B.cv + 3 * size_t(e)) + dv);
                            tst(T.vel, tld(T.vel) - tool_linear(T, nv));
                            tst(T.ang, tld(T.ang) - tool_angular(T, tcross(arm, nv)));
                            const t3 after = tld(T.vel) + tcross(tld(T.ang), arm) - vi;
                            const t3 tangent = after - nrm * tdot(after, nrm);
                            const double ts = tlen(tangent);
                            const double rcp = 1.0 / ts;
                            const t3 tdir = (isfinite(rcp) && rcp > 0.0) ? tangent * rcp : tv(0, 0, 0);
                            const double tw = !kin ? tdot(thmul(tdir, fr), tdir) / mi : 0.0;
                            const double tden = tw + tdot(tdir, tool_linear(T, tdir)) +
                                                tdot(tdir, tcross(tool_angular(T, tcross(arm, tdir)), arm));
                            if (ts > 1e-6 && tden > 1e-9) {
                                const double fm = fmin(ts / tden, T.friction * ni);
                                const t3 f = tdir * fm;
                                const t3 fdv = tdiv(thmul(f, fr), mi);
                                vi = vi + fdv;
                                if (timed) tst(B.cv + 3 * size_t(e), tld(B.cv + 3 * size_t(e)) + fdv);
                                tst(T.vel, tld(T.vel) - tool_linear(T, f));
                                tst(T.ang, tld(T.ang) - tool_angular(T, tcross(arm, f)));
                                T.friction_heat += fmax(fm * ts - 0.5 * fm * fm * tden, 0.0);
                            }
                            A.vel[v] = vi.x;
                            A.vel[R + v] = vi.y;
                            A.vel[2 * R + v] = vi.z;
                        }
                        if (tc.pen > max_pen) {
                            max_pen = tc.pen;
                            correction = tool_correction(T, nrm, tc.pen);
                        }
                    }
                }
                tst(T.pos, tld(T.pos) + correction);
                tool_constrain(T);
                if (tool_moved(T) > T.reach) escaped = 1;
            }
            __syncthreads();
            if (escaped) break;
            // Re-projection until no overlap deeper than 1e-6 is left (at most
            // six passes): the first deepest overlap in the reference's order.
            for (int pass = 0; pass < 6; ++pass) {
                double my_pen = 1e-6;
                uint32_t my_cell = ~0u;
                const tq rp = tqld(T.rot);
                for (uint32_t c = threadIdx.x; c < nc; c += kToolBlock) {
                    const t3 center = tld(T.pos) + trot(rp, tld(B.cells + 3 * c));
                    double cp = 1e-6;
                    t3 cn{0, 0, 0};
                    tool_cell_pairs(B, C, center, axes, [&](uint32_t, uint32_t v) {
                        t3 nrm;
                        double pen;
                        if (!tool_sat(center, axes, arena_pos(A, v), nrm, pen)) return;
                        if (pen > cp) {
                            cp = pen;
                            cn = nrm;
                        }
                    });
                    tst(B.con[size_t(c) * kToolCellSlots].n, cn);
                    if (cp > my_pen) {  // cells ascend per thread: the first deepest
                        my_pen = cp;
                        my_cell = c;
                    }
                }
                // Workgroup reduction: the deepest, the lowest cell among equals.
                for (int off = 32; off > 0; off >>= 1) {
                    const double op = __shfl_down(my_pen, off);
                    const uint32_t oc = __shfl_down(my_cell, off);
                    if (op > my_pen || (op == my_pen && oc < my_cell)) {
                        my_pen = op;
                        my_cell = oc;
                    }
                }
                if ((threadIdx.x & 63u) == 0) {
                    red_pen[threadIdx.x >> 6] = my_pen;
                    red_cell[threadIdx.x >> 6] = my_cell;
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    double bp = red_pen[0];
                    uint32_t bc = red_cell[0];
                    for (uint32_t w = 1; w < kToolBlock / 64; ++w)
                        if (red_pen[w] > bp || (red_pen[w] == bp && red_cell[w] < bc)) {
                            bp = red_pen[w];
                            bc = red_cell[w];
                        }
                    done = bp <= 1e-6 ? 1u : 0u;
                    if (!done) {
                        tst(T.pos, tld(T.pos) + tool_correction(T, tld(B.con[size_t(bc) * kToolCellSlots].n), bp));
                        tool_constrain(T);
                        if (tool_moved(T) > T.reach) escaped = 1;
                    }
                }
                __syncthreads();
                if (done || escaped) break;
            }
            if (escaped) break;
        }
        if (!escaped) break;
        // Undo the attempt: the REV velocities it changed and the tool state.
        for (uint32_t e = threadIdx.x; e < C; e += kToolBlock) {
            const uint32_t v = B.idx2[e];
            for (int q = 0; q < 3; ++q) A.vel[q * R + v] = B.vsave[3 * size_t(e) + q];
        }
        __syncthreads();  // every thread has read `escaped` before it is cleared
        if (threadIdx.x == 0) {
            const double reach = T.reach * 2.0;
            T = T0;
            T.reach = reach;
            const double pad = reach + 4.2 * kToolCell;
            for (int k = 0; k < 3; ++k) {
                T.box_lo[k] = T.cell_lo[k] - pad;
                T.box_hi[k] = T.cell_hi[k] + pad;
            }
            escaped = 0;
            cand = 0;
            if (attempt + 1 == kToolAttempts) lstatus |= TS_MARGIN;
        }
        __syncthreads();
        if (lstatus & TS_MARGIN) break;
        // The candidates of the larger box, in index order.
        for (uint32_t base = 0; base < g.n; base += kToolBlock) {
            const uint32_t local = base + threadIdx.x;
            bool in = false;
            t3 p{0, 0, 0};
            if (local < g.n) {
                p = arena_pos(A, g.voff + local);
                in = tool_in_box(T, p);
            }
            const uint64_t m = __ballot(in);
            const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
            if (lane == 0) wcnt[w] = uint32_t(__popcll(m));
            __syncthreads();
            uint32_t at = cand + uint32_t(__popcll(m & ((uint64_t(1) << lane) - 1)));
            for (uint32_t k = 0; k < w; ++k) at += wcnt[k];
            if (in && at < kToolMaxCand) {
                int32_t k3[3];
                uint64_t key = ~uint64_t(0);
                if (tool_key(p, k3))
                    key = tool_pack(k3[0], k3[1], k3[2]);
                else
                    atomicOr(&lstatus, uint32_t(TS_KEY));
                B.key[at] = key;
                B.idx[at] = g.voff + local;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t add = 0;
                for (uint32_t k = 0; k < kToolBlock / 64; ++k) add += wcnt[k];
                cand += add;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0 && cand > kToolMaxCand) lstatus |= TS_CAND;
        __syncthreads();
        if (lstatus & (TS_CAND | TS_KEY)) break;
    }
    // Contact velocity: half a substep of drift, and the final velocities
    // (the zero-time pass discards it).
    const bool failed = lstatus & (TS_MARGIN | TS_CAND | TS_KEY);
    for (uint32_t e = threadIdx.x; timed && !failed && e < cand; e += kToolBlock) {
        const t3 cv = tld(B.cv + 3 * size_t(e));
        if (cv.x == 0.0 && cv.y == 0.0 && cv.z == 0.0) continue;
        const uint32_t v = B.idx2[e];
        const t3 p = arena_pos(A, v) + cv * (0.5 * dt);
        A.pos[v] = p.x;
        A.pos[R + v] = p.y;
        A.pos[2 * R + v] = p.z;
        A.pred[v] = p.x;
        A.pred[R + v] = p.y;
        A.pred[2 * R + v] = p.z;
        A.drift[v] += cv.x;
        A.drift[R + v] += cv.y;
        A.drift[2 * R + v] += cv.z;
    }
    if (threadIdx.x == 0) {
        T.contacts = T.contacts + contacts;
        T.candidates = cand;
        T.status = T.status | lstatus;
        *B.dev = T;
    }
}

// One workgroup per scene: the fracture pass's totals over the scene's tiles
// (fixed order) into its state.
__global__ __launch_bounds__(kBlock) void reduce_fracture(DevArena A) {
    const uint32_t slot = blockIdx.x;
    if (!(A.bits[slot] & B_FRACTURE)) return;
    const SlotGeo g = A.geo[slot];
    if (g.n == 0) return;
    __shared__ double lds[kFracN][kWaves];
    double acc[kFracN] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < g.ntiles; b += kBlock)
        for (int q = 0; q < kFracN; ++q) acc[q] += A.partials[size_t(kFrac + q) * A.cap_tiles + g.tile0 + b];
    block_sum_all<kFracN>(acc, lds);
    if (threadIdx.x == 0)
        for (int q = 0; q < kFracN; ++q) A.st[slot].frac[q] = acc[q];
}

// The scenes' states into pinned host memory at the end of a tick; the host
// reads them after the tick's event.
__global__ __launch_bounds__(kBlock) void publish_state(const SlotState *__restrict__ src, SlotState *dst, uint32_t n) {
    const uint64_t *s = reinterpret_cast<const uint64_t *>(src);
    uint64_t *d = reinterpret_cast<uint64_t *>(dst);
    const uint32_t words = n * uint32_t(sizeof(SlotState) / 8);
    for (uint32_t i = threadIdx.x; i < words; i += kBlock) d[i] = s[i];
}

// Velocities after a substep (src/solver.rs:668-683), no contact.
__global__ __launch_bounds__(kBlock) void finish_velocities(DevArena A) {
    Tile c;
    if (!tile_for(A, B_FINISH, c) || !c.row) return;
    const size_t R = A.R;
    const uint32_t v = c.v;
    if (A.kinematic[v]) return;
    const double h = A.prm[c.slot].h;
    const uint8_t tm = A.tmask[v];
    for (int q = 0; q < 3; ++q) {
        const double p = A.pos[q * R + v];
        if (tm >> q & 1)
            A.vel[q * R + v] = (p - A.prev[q * R + v]) / h;
        else
            A.vel[q * R + v] = A.drift[q * R + v] + (p - A.pred[q * R + v]) / h;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
inline uint32_t tiles_for(uint64_t rows) { return uint32_t((rows + kBlock - 1) / kBlock); }

// HBM bytes per arena row: the planes carved in create_batch.
constexpr size_t kRowBytes = 24 * 5 + 32 + 24 * 3 + 24 + 4 * 9 + 12 + 12 + 4 + 12 + 12 + 36 + 36 + 144 + 72 + 48 + 48 + 24 +
                             72 + 24 + 24 + 6 * 12 + 6 * 24 + 4 * kStrength + 12 * kHistory + 32 + 72;

enum Phase : uint32_t { P_IDLE = 0, P_START, P_SOLVE, P_APPLIED, P_FINISHED, P_DONE, P_FAILED };
constexpr uint32_t kNoSeg = 0xFFFFFFFFu;

struct SlotHost {
    bool used = false;
    SlotGeo geo{};
    uint32_t active = 0;
    vx_step_params params{};
    uint32_t steps = 0, steps_done = 0, substep = 0, pass = 0;
    Phase phase = P_IDLE;
    bool warm_ok = false;
    bool finish_next = false;
    uint32_t status = VX_OK;
    uint32_t solves = 0;
    uint64_t inner = 0, sweeps = 0;
    double last_rel = 0.0, last_angle = 0.0;
    bool manual = false;  // driven by the vx_scene_* stage calls
    bool held = false;    // sits out the ticks until released (vx_batch_hold)
    // The last inner residual seen in the current segment (for the chunk
    // length of a one-scene arena): iteration, r.r, segment start.
    uint32_t seen_k = 0, seen_seg = kNoSeg;
    double seen_rr = 0.0;
    bool has_strength = false;
    bool has_tool = false;      // a vx_scene with a rigid tool (vx_scene_set_tool)
    vx_fracture_report frac{};  // fracture totals since admission (a vx_scene: this step)
};

// The device buffers of a one-scene arena's rigid tool.
struct ToolHost {
    ToolBufs bufs{};
    void *alloc = nullptr;
    uint32_t n_cells = 0;
};

// One launched tick: the bits it ran with and the scene states read back
// after it (both pinned), and the event that marks its end.
struct Flight {
    uint32_t *bits = nullptr;
    SlotState *st = nullptr;      // pinned, coherent: written by publish_state
    SlotState *st_dev = nullptr;  // its device address
    hipEvent_t done = nullptr;
    hipEvent_t frac0 = nullptr, frac1 = nullptr;  // around the fracture kernels
    uint32_t *list = nullptr;                     // pinned: the tick's tile list (when it has one)
    uint32_t iters = 0;                           // inner iterations of its chunk
    bool active = false;
    bool frac_timed = false;
};

// ---- aggregation multigrid preconditioner ------------------------------------
// One V(1,1) cycle of aggregation multigrid as the preconditioner of the
// FP32 inner PCG of a one-scene arena (VX_PRECOND_MULTIGRID). Level l + 1
// aggregates the nodes of level l in 2x2x2 boxes (the last box along an axis
// of odd extent holds one layer): the prolongation P copies a coarse node's
// value to its children, component by component, on the components the child
// leaves free, and the coarse operator is the Galerkin product P^T A P. With
// the bond stencil that product keeps the stencil's shape: per coarse node a
// symmetric diagonal block D (the children's free diagonal blocks, less the
// couplings between children in both directions) and one 3x3 coupling C per
// axis to the next node along it, A[I, I + e] = -C and A[I + e, I] = -C^T.
// A component no child leaves free is fixed on the coarse level too (identity
// diagonal, no coupling). The fine level is the scene's own operator: D and
// the bond blocks k with the free-component masks, A[v, v + e] = -F_v k F_j.
// Smoothing is damped block-Jacobi, x += w D^-1 (r - A x), once before
// (from x = 0: x = w D^-1 r) and once after the coarse correction; the
// coarsest level (at most kMgCoarsest nodes) is solved exactly, e = A^-1 r,
// with the inverse found once per assembly in FP64 from its Cholesky factor
// (A^-1 = W^T W, W = L^-1). The hierarchy is built in FP64 from the FP64 operator after every
// assembly and applied in FP32. The cycle is a fixed symmetric positive
// definite linear operator, so the outer iteration stays conjugate gradients.
// Every level's matrix graph is the 7-point lattice, which two colours
// separate, so the eigenvalues of D^-1 A lie in (0, 2): any w < 1 keeps the
// smoother convergent and the cycle positive definite. Measured on the 64^3
// system: w = 0.5 / 0.6 / 0.7 / 0.8 / 0.9 take 63 / 60 / 58 / 56 / 55
// iterations.
constexpr float kMgOmega = 0.9f;
constexpr uint32_t kMgCoarsest = 64;           // nodes of the level solved directly, at most
constexpr uint32_t kMgMaxDense = 3 * kMgCoarsest;
constexpr uint32_t kMgMaxLevels = 16;
constexpr int kMgBlock = 256;
constexpr int kMgDenseBlock = 512;             // one workgroup for the dense factor and its use
constexpr uint32_t kMgChunk = 4;               // inner iterations per chunk at most

struct MgLevel {
    uint32_t nx, ny, nz, n;  // node grid of the level
    float *D, *C, *Dinv;     // 6n symmetric, 27n ([axis][3x3 row-major]), 6n symmetric
    float *r, *x, *res;      // 3n each: right-hand side, pre-smoothed iterate, residual (after the cycle: the level's result)
    uint8_t *mask;           // free components
    double *D64, *C64;       // FP64 copies of D and C (the next level is built from them)
};

__device__ __forceinline__ int sym_at(int r, int c) {
    return r <= c ? (r == 0 ? c : (r == 1 ? 2 + c : 5)) : (c == 0 ? r : (c == 1 ? 2 + r : 5));
}
__device__ __forceinline__ void node_xyz(uint32_t i, uint32_t nx, uint32_t ny, uint32_t (&c)[3]) {
    c[0] = i % nx;
    const uint32_t t = i / nx;
    c[1] = t % ny;
    c[2] = t / ny;
}
__device__ __forceinline__ uint32_t node_stride(uint32_t nx, uint32_t ny, int a) { return a == 0 ? 1u : (a == 1 ? nx : nx * ny); }

// Store a level's node from its FP64 diagonal block and couplings.
__device__ __forceinline__ void mg_store_node(const MgLevel &L, uint32_t I, double (&D)[3][3], const double (&C)[3][3][3],
                                              uint32_t fm) {
    for (int r = 0; r < 3; ++r)
        if (!(fm >> r & 1u)) D[r][r] = 1.0;  // its row and column are zero: identity
    double s[6] = {D[0][0], D[0][1], D[0][2], D[1][1], D[1][2], D[2][2]}, si[6];
    sym_inverse(s, si);
    for (int q = 0; q < 6; ++q) {
        L.D64[size_t(q) * L.n + I] = s[q];
        L.D[size_t(q) * L.n + I] = float(s[q]);
        L.Dinv[size_t(q) * L.n + I] = float(si[q]);
    }
    for (int a = 0; a < 3; ++a)
        for (int e = 0; e < 9; ++e) {
            L.C64[size_t(a * 9 + e) * L.n + I] = C[a][e / 3][e % 3];
            L.C[size_t(a * 9 + e) * L.n + I] = float(C[a][e / 3][e % 3]);
        }
    L.mask[I] = uint8_t(fm);
}

// Level 1 from the scene's FP64 operator: one thread per coarse node.
__global__ __launch_bounds__(kMgBlock) void mg_build_fine(DevArena A, SlotGeo g, MgLevel L) {
    const uint32_t I = blockIdx.x * kMgBlock + threadIdx.x;
    if (I >= L.n) return;
    uint32_t ci[3];
    node_xyz(I, L.nx, L.ny, ci);
    const size_t R = A.R;
    const uint32_t ext[3] = {g.nx, g.ny, g.nz};
    double D[3][3] = {}, C[3][3][3] = {};
    uint32_t fm = 0;
    for (int d = 0; d < 8; ++d) {
        const uint32_t off[3] = {uint32_t(d & 1), uint32_t(d >> 1 & 1), uint32_t(d >> 2)};
        const uint32_t c[3] = {2 * ci[0] + off[0], 2 * ci[1] + off[1], 2 * ci[2] + off[2]};
        if (c[0] >= ext[0] || c[1] >= ext[1] || c[2] >= ext[2]) continue;
        const uint32_t v = g.voff + c[0] + g.nx * (c[1] + g.ny * c[2]);
        const uint32_t m = A.mask[v];
        fm |= m;
        const double f[3] = {double(m & 1u), double(m >> 1 & 1u), double(m >> 2 & 1u)};
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) D[r][q] += f[r] * A.d64[size_t(sym_at(r, q)) * R + v] * f[q];
        for (int a = 0; a < 3; ++a) {
            if (c[a] + 1 >= ext[a]) continue;  // no +a neighbour
            const uint32_t mj = A.mask[v + stride_of(g, a)];
            const double fj[3] = {double(mj & 1u), double(mj >> 1 & 1u), double(mj >> 2 & 1u)};
            double k[3][3];
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) k[r][q] = f[r] * A.k64[(size_t(a) * 6 + sym_at(r, q)) * R + v] * fj[q];
            if (off[a] == 0) {  // the neighbour is a child of the same node
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) D[r][q] -= k[r][q] + k[q][r];
            } else {  // it is a child of the next node along a
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) C[a][r][q] += k[r][q];
            }
        }
    }
    mg_store_node(L, I, D, C, fm);
}

// Level l + 1 from level l (S).
__global__ __launch_bounds__(kMgBlock) void mg_build_coarse(MgLevel S, MgLevel L) {
    const uint32_t I = blockIdx.x * kMgBlock + threadIdx.x;
    if (I >= L.n) return;
    uint32_t ci[3];
    node_xyz(I, L.nx, L.ny, ci);
    const uint32_t ext[3] = {S.nx, S.ny, S.nz};
    double D[3][3] = {}, C[3][3][3] = {};
    uint32_t fm = 0;
    for (int d = 0; d < 8; ++d) {
        const uint32_t off[3] = {uint32_t(d & 1), uint32_t(d >> 1 & 1), uint32_t(d >> 2)};
        const uint32_t c[3] = {2 * ci[0] + off[0], 2 * ci[1] + off[1], 2 * ci[2] + off[2]};
        if (c[0] >= ext[0] || c[1] >= ext[1] || c[2] >= ext[2]) continue;
        const uint32_t i = c[0] + S.nx * (c[1] + S.ny * c[2]);
        const uint32_t m = S.mask[i];
        fm |= m;
        const double f[3] = {double(m & 1u), double(m >> 1 & 1u), double(m >> 2 & 1u)};
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) D[r][q] += f[r] * S.D64[size_t(sym_at(r, q)) * S.n + i] * f[q];
        for (int a = 0; a < 3; ++a) {
            if (c[a] + 1 >= ext[a]) continue;
            double k[3][3];  // S's couplings carry their masks already
            for (int e = 0; e < 9; ++e) k[e / 3][e % 3] = S.C64[size_t(a * 9 + e) * S.n + i];
            if (off[a] == 0) {
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) D[r][q] -= k[r][q] + k[q][r];
            } else {
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) C[a][r][q] += k[r][q];
            }
        }
    }
    mg_store_node(L, I, D, C, fm);
}

// The coarsest level as a dense n3 x n3 matrix (n3 = 3n, row 3I + r): each
// thread writes the three rows of its node.
__global__ __launch_bounds__(kMgBlock) void mg_dense(MgLevel L, double *M, uint32_t n3) {
    const uint32_t I = blockIdx.x * kMgBlock + threadIdx.x;
    if (I >= L.n) return;
    uint32_t ci[3];
    node_xyz(I, L.nx, L.ny, ci);
    const uint32_t ext[3] = {L.nx, L.ny, L.nz};
    for (int r = 0; r < 3; ++r) {
        double *row = M + size_t(3 * I + r) * n3;
        for (uint32_t j = 0; j < n3; ++j) row[j] = 0.0;
        for (int q = 0; q < 3; ++q) row[3 * I + q] = L.D64[size_t(sym_at(r, q)) * L.n + I];
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = node_stride(L.nx, L.ny, a);
            if (ci[a] + 1 < ext[a])
                for (int q = 0; q < 3; ++q) row[3 * (I + s) + q] = -L.C64[size_t(a * 9 + r * 3 + q) * L.n + I];
            if (ci[a] > 0)
                for (int q = 0; q < 3; ++q) row[3 * (I - s) + q] = -L.C64[size_t(a * 9 + q * 3 + r) * L.n + I - s];
        }
    }
}

// Cholesky factor of M (in LDS, the lower triangle packed by rows: at most
// 148 KB for 192 unknowns), then W = L^-1 (lower), stored in FP32 (W^T W
// stays symmetric positive definite whatever W's rounding). M is banded
// (|row - column| <= bw: the stencil's reach in the lexicographic node
// order), and so is L. One workgroup factors it (right-looking: at step k
// the column below the pivot is scaled and the trailing band updated by
// every thread); L then goes to global memory (over M, packed) and the same LDS
// holds W while it is found row by row by forward substitution (one column
// per thread, FP64 sums, the band rows of L staged in LDS); finally M
// receives the inverse W^T W in FP64. A non-positive
// pivot sets *err and is replaced by 1 (the cycle stays defined).
__device__ __forceinline__ uint32_t tri(uint32_t i, uint32_t j) { return i * (i + 1) / 2 + j; }
constexpr uint32_t kMgLds = kMgMaxDense * (kMgMaxDense + 1) / 2 * 8;  // bytes: packed FP64 L, or padded FP32 W
static_assert(kMgMaxDense * (kMgMaxDense + 1) * 4 <= kMgLds, "W fits where L was");
constexpr uint32_t kMgLoads = 8;  // loads in flight per thread when moving the dense blocks
constexpr uint32_t kMgRowBlock = 8;  // rows of L staged together while W is found
__global__ __launch_bounds__(kMgDenseBlock) void mg_cholesky(double *M, uint32_t n3, uint32_t bw, uint32_t *err) {
    __shared__ double lds[kMgLds / 8];
    __shared__ double lblk[kMgRowBlock * kMgMaxDense];
    double *Lp = lds;
    const uint32_t t = threadIdx.x, nn = n3 * n3;
    for (uint32_t e0 = t; e0 < nn; e0 += kMgDenseBlock * kMgLoads) {  // the lower triangle of M
        double v[kMgLoads];
#pragma unroll
        for (uint32_t u = 0; u < kMgLoads; ++u) {
            const uint32_t e = e0 + u * kMgDenseBlock;
            v[u] = e < nn ? M[e] : 0.0;
        }
#pragma unroll
        for (uint32_t u = 0; u < kMgLoads; ++u) {
            const uint32_t e = e0 + u * kMgDenseBlock, i = e / n3, j = e % n3;
            if (e < nn && j <= i) Lp[tri(i, j)] = v[u];
        }
    }
    for (uint32_t k = 0; k < n3; ++k) {
        __syncthreads();  // the triangle is in LDS; the previous step's updates are done
        const double piv = Lp[tri(k, k)];
        const double d = piv > 0.0 ? sqrt(piv) : 1.0;
        if (!(piv > 0.0) && t == 0) *err = 1u;
        const uint32_t end = min(n3, k + bw + 1);  // rows below k with entries in column k
        for (uint32_t i = k + 1 + t; i < end; i += kMgDenseBlock) Lp[tri(i, k)] /= d;
        __syncthreads();  // column k is scaled; every thread has read the pivot
        if (t == 0) Lp[tri(k, k)] = d;
        // the trailing band, (i, j) with k < j <= i < end, spread over every thread
        const uint32_t nb = end - k - 1, ne = nb * (nb + 1) / 2;
        for (uint32_t e = t; e < ne; e += kMgDenseBlock) {
            uint32_t ii = uint32_t((sqrtf(8.0f * float(e) + 1.0f) - 1.0f) * 0.5f);
            while (ii * (ii + 1) / 2 > e) --ii;  // exact for any rounding of the root
            while ((ii + 1) * (ii + 2) / 2 <= e) ++ii;
            const uint32_t i = k + 1 + ii, j = k + 1 + (e - ii * (ii + 1) / 2);
            Lp[tri(i, j)] -= Lp[tri(i, k)] * Lp[tri(j, k)];
        }
    }
    __syncthreads();
    const uint32_t nt = n3 * (n3 + 1) / 2;
    for (uint32_t e = t; e < nt; e += kMgDenseBlock) M[e] = Lp[e];  // L, packed
    __syncthreads();  // L is in global memory and every thread is done with Lp
    float *Wl = reinterpret_cast<float *>(lds);  // W[k][c] at k * (n3 + 1) + c
    const uint32_t ld = n3 + 1;
    for (uint32_t e = t; e < n3 * ld; e += kMgDenseBlock) Wl[e] = 0.0f;
    __syncthreads();
    // W row by row: W[k][c] = (e_k - sum_{m<k} L[k][m] W[m][c])_c / L[k][k] for
    // c <= k, one column per thread; L[k][m] is zero for m < k - bw. The band
    // rows of L are staged in LDS eight at a time.
    for (uint32_t k0 = 0; k0 < n3; k0 += kMgRowBlock) {
        for (uint32_t r = 0; r < kMgRowBlock && k0 + r < n3; ++r) {
            const uint32_t k = k0 + r, m0 = k > bw ? k - bw : 0u;
            for (uint32_t m = m0 + t; m <= k; m += kMgDenseBlock) lblk[r * kMgMaxDense + (m - m0)] = M[tri(k, m)];
        }
        __syncthreads();
        for (uint32_t r = 0; r < kMgRowBlock && k0 + r < n3; ++r) {
            const uint32_t k = k0 + r, m0 = k > bw ? k - bw : 0u;
            const double *lr = lblk + r * kMgMaxDense;  // lr[m - m0] = L[k][m]
            const double lkk = lr[k - m0];
            for (uint32_t c = t; c <= k; c += kMgDenseBlock) {
                double s = c == k ? 1.0 : 0.0;
                for (uint32_t m = max(c, m0); m < k; ++m) s -= lr[m - m0] * double(Wl[m * ld + c]);
                Wl[k * ld + c] = float(s / lkk);
            }
            __syncthreads();  // row k is complete before the next rows read it (and lblk before its reload)
        }
    }
    // the inverse of the coarsest matrix, A^-1 = W^T W (FP64, over M)
    for (uint32_t e = t; e < nn; e += kMgDenseBlock) {
        const uint32_t a = e / n3, b = e % n3;
        double acc = 0.0;
        for (uint32_t k = max(a, b); k < n3; ++k) acc = __builtin_fma(double(Wl[k * ld + a]), double(Wl[k * ld + b]), acc);
        M[e] = acc;
    }
}

// x0 = w D^-1 r of fine row v: the update kernel (or the residual
// replacement) before the cycle leaves z = D^-1 r (zero on fixed components:
// r is zero there and D^-1 is the identity).
__device__ __forceinline__ void mg_x0(const DevArena &A, uint32_t v, float (&x)[3]) {
    const size_t R = A.R;
    for (int q = 0; q < 3; ++q) x[q] = kMgOmega * A.z[q * R + v];
}
__device__ __forceinline__ void sym_mul_acc(const float (&s)[6], const float (&x)[3], float (&y)[3]) {
    y[0] = __builtin_fmaf(s[2], x[2], __builtin_fmaf(s[1], x[1], __builtin_fmaf(s[0], x[0], y[0])));
    y[1] = __builtin_fmaf(s[4], x[2], __builtin_fmaf(s[3], x[1], __builtin_fmaf(s[1], x[0], y[1])));
    y[2] = __builtin_fmaf(s[5], x[2], __builtin_fmaf(s[4], x[1], __builtin_fmaf(s[2], x[0], y[2])));
}

// Fine level, before the coarse correction: res = r - A x0 with x0 = w D^-1 r.
// NT (arenas beyond the Infinity Cache): D and the stores stream.
template <bool NT>
__global__ __launch_bounds__(kMgBlock) void mg_fine_down(DevArena A, SlotGeo g, const SlotState *st, float *res,
                                                         float *x0) {
    const uint32_t loc = blockIdx.x * kMgBlock + threadIdx.x;
    if (st->done || loc >= g.n) return;
    const size_t R = A.R;
    const uint32_t v = g.voff + loc;
    uint32_t cc[3];
    coords(g, loc, cc);
    float x[3], d[6], y[3] = {0.0f, 0.0f, 0.0f}, acc[3] = {0.0f, 0.0f, 0.0f};
    mg_x0(A, v, x);
    for (int q = 0; q < 6; ++q) d[q] = NT ? __builtin_nontemporal_load(A.d32 + q * R + v) : A.d32[q * R + v];
    sym_mul_acc(d, x, y);
    for (int a = 0; a < 3; ++a) {
        const uint32_t s = stride_of(g, a);
        float xn[3], k[6];
        if (cc[a] + 1 < extent_of(g, a)) {
            mg_x0(A, v + s, xn);
            for (int q = 0; q < 6; ++q) k[q] = A.k32[(size_t(a) * 6 + q) * R + v];
            sym_mul_acc(k, xn, acc);
        }
        if (cc[a] > 0) {
            mg_x0(A, v - s, xn);
            for (int q = 0; q < 6; ++q) k[q] = A.k32[(size_t(a) * 6 + q) * R + v - s];
            sym_mul_acc(k, xn, acc);
        }
    }
    const uint32_t m = A.mask[v];
    for (int q = 0; q < 3; ++q) {
        const float rq = A.r[q * R + v] - (y[q] - ((m >> q & 1u) ? acc[q] : 0.0f));
        if constexpr (NT) {
            __builtin_nontemporal_store(rq, res + size_t(q) * g.n + loc);
            __builtin_nontemporal_store(x[q], x0 + size_t(q) * g.n + loc);
        } else {
            res[size_t(q) * g.n + loc] = rq;
            x0[size_t(q) * g.n + loc] = x[q];  // for the fine level's post-smoothing (z is overwritten there)
        }
    }
}

// Restriction to level L (children from the fine level or from level S) and
// the pre-smoothing there: r_I = sum of the children's free residuals,
// x_I = w D_I^-1 r_I.
__device__ __forceinline__ void mg_restrict_store(const MgLevel &L, uint32_t I, const float (&r)[3]) {
    float di[6], x[3] = {0.0f, 0.0f, 0.0f};
    for (int q = 0; q < 6; ++q) di[q] = L.Dinv[size_t(q) * L.n + I];
    sym_mul_acc(di, r, x);
    for (int q = 0; q < 3; ++q) {
        L.r[size_t(q) * L.n + I] = r[q];
        L.x[size_t(q) * L.n + I] = kMgOmega * x[q];
    }
}
__global__ __launch_bounds__(kMgBlock) void mg_restrict_fine(DevArena A, SlotGeo g, const SlotState *st,
                                                             const float *res, MgLevel L) {
    const uint32_t I = blockIdx.x * kMgBlock + threadIdx.x;
    if (st->done || I >= L.n) return;
    uint32_t ci[3];
    node_xyz(I, L.nx, L.ny, ci);
    float r[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < 8; ++d) {
        const uint32_t c[3] = {2 * ci[0] + (d & 1), 2 * ci[1] + (d >> 1 & 1), 2 * ci[2] + (d >> 2)};
        if (c[0] >= g.nx || c[1] >= g.ny || c[2] >= g.nz) continue;
        const uint32_t loc = c[0] + g.nx * (c[1] + g.ny * c[2]);
        const uint32_t m = A.mask[g.voff + loc];
        for (int q = 0; q < 3; ++q)
            if (m >> q & 1u) r[q] += res[size_t(q) * g.n + loc];
    }
    mg_restrict_store(L, I, r);
}
__global__ __launch_bounds__(kMgBlock) void mg_restrict_coarse(MgLevel S, MgLevel L, const SlotState *st) {
    const uint32_t I = blockIdx.x * kMgBlock + threadIdx.x;
    if (st->done || I >= L.n) return;
    uint32_t ci[3];
    node_xyz(I, L.nx, L.ny, ci);
    float r[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < 8; ++d) {
        const uint32_t c[3] = {2 * ci[0] + (d & 1), 2 * ci[1] + (d >> 1 & 1), 2 * ci[2] + (d >> 2)};
        if (c[0] >= S.nx || c[1] >= S.ny || c[2] >= S.nz) continue;
        const uint32_t i = c[0] + S.nx * (c[1] + S.ny * c[2]);
        const uint32_t m = S.mask[i];
        for (int q = 0; q < 3; ++q)
            if (m >> q & 1u) r[q] += S.res[size_t(q) * S.n + i];
    }
    mg_restrict_store(L, I, r);
}

// y = A_L x at node I of a coarse level, with x given per node by `xat`.
template <class X>
__device__ __forceinline__ void mg_coarse_apply(const MgLevel &L, uint32_t I, const X &xat, float (&y)[3]) {
    uint32_t ci[3];
    node_xyz(I, L.nx, L.ny, ci);
    const uint32_t ext[3] = {L.nx, L.ny, L.nz};
    float x[3], d[6];
    xat(I, x);
    for (int q = 0; q < 6; ++q) d[q] = L.D[size_t(q) * L.n + I];
    y[0] = y[1] = y[2] = 0.0f;
    sym_mul_acc(d, x, y);
    for (int a = 0; a < 3; ++a) {
        const uint32_t s = node_stride(L.nx, L.ny, a);
        float xn[3];
        if (ci[a] + 1 < ext[a]) {  // -C x_{I+e}
            xat(I + s, xn);
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) y[r] -= L.C[size_t(a * 9 + r * 3 + q) * L.n + I] * xn[q];
        }
        if (ci[a] > 0) {  // -C_{I-e}^T x_{I-e}
            xat(I - s, xn);
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) y[r] -= L.C[size_t(a * 9 + q * 3 + r) * L.n + I - s] * xn[q];
        }
    }
}

// Coarse level, before the correction from below: res = r - A x.
__global__ __launch_bounds__(kMgBlock) void mg_coarse_down(MgLevel L, const SlotState *st) {
    const uint32_t I = blockIdx.x * kMgBlock + threadIdx.x;
    if (st->done || I >= L.n) return;
    auto xat = [&](uint32_t u, float (&x)[3]) {
        for (int q = 0; q < 3; ++q) x[q] = L.x[size_t(q) * L.n + u];
    };
    float y[3];
    mg_coarse_apply(L, I, xat, y);
    for (int q = 0; q < 3; ++q) L.res[size_t(q) * L.n + I] = L.r[size_t(q) * L.n + I] - y[q];
}

// The coarsest level: res = A^-1 r with the dense inverse found at set-up;
// one wave per row, its lanes splitting the row (the matrix is symmetric).
__global__ __launch_bounds__(kMgDenseBlock) void mg_coarsest(MgLevel L, const double *Ainv, uint32_t n3, const SlotState *st) {
    if (st->done) return;
    const uint32_t lane = threadIdx.x & 63u, i = blockIdx.x * (kMgDenseBlock / 64) + (threadIdx.x >> 6);
    if (i >= n3) return;  // the whole wave
    double s = 0.0;
    for (uint32_t m = lane; m < n3; m += 64)
        s = __builtin_fma(Ainv[size_t(i) * n3 + m], double(L.r[size_t(m % 3) * L.n + m / 3]), s);
    s = wave_sum(s);
    if (lane == 0) L.res[size_t(i % 3) * L.n + i / 3] = float(s);
}

// Coarse level, after the correction: x' = x + P e (e: the coarser level's
// result, in P.res), then x'' = x' + w D^-1 (r - A x'), stored in L.res (the
// level's result).
__global__ __launch_bounds__(kMgBlock) void mg_coarse_up(MgLevel L, MgLevel P, const SlotState *st) {
    const uint32_t I = blockIdx.x * kMgBlock + threadIdx.x;
    if (st->done || I >= L.n) return;
    auto xat = [&](uint32_t u, float (&x)[3]) {
        uint32_t c[3];
        node_xyz(u, L.nx, L.ny, c);
        const uint32_t pu = c[0] / 2 + P.nx * (c[1] / 2 + P.ny * (c[2] / 2));
        const uint32_t m = L.mask[u];
        for (int q = 0; q < 3; ++q)
            x[q] = L.x[size_t(q) * L.n + u] + ((m >> q & 1u) ? P.res[size_t(q) * P.n + pu] : 0.0f);
    };
    float y[3], x[3], di[6], dx[3] = {0.0f, 0.0f, 0.0f}, rr[3];
    mg_coarse_apply(L, I, xat, y);
    xat(I, x);
    for (int q = 0; q < 6; ++q) di[q] = L.Dinv[size_t(q) * L.n + I];
    for (int q = 0; q < 3; ++q) rr[q] = L.r[size_t(q) * L.n + I] - y[q];
    sym_mul_acc(di, rr, dx);
    for (int q = 0; q < 3; ++q) L.res[size_t(q) * L.n + I] = __builtin_fmaf(kMgOmega, dx[q], x[q]);
}

// Fine level, after the correction: z = x' + w D^-1 (r - A x') with
// x' = x0 + P e, and the tile's partial of r.z (the preconditioned residual
// the next SpMV's head takes). Tiles as the inner kernels of the scene.
// NT: D, D^-1 and the store of z stream.
template <bool NT>
__global__ __launch_bounds__(kBlock) void mg_fine_up(DevArena A, OneScene one, MgLevel P, const float *x0) {
    Tile c;
    if (!inner_tile<float, true>(A, one, c)) return;
    __shared__ double lds[1][kWaves];
    if (A.st[c.slot].done) return;
    const size_t R = A.R;
    const SlotGeo &g = c.g;
    double part[1] = {0.0};
    if (c.row) {
        const uint32_t v = c.v;
        auto xat = [&](uint32_t u, float (&x)[3]) {
            uint32_t cu[3];
            coords(g, u - g.voff, cu);
            const uint32_t pu = cu[0] / 2 + P.nx * (cu[1] / 2 + P.ny * (cu[2] / 2));
            const uint32_t m = A.mask[u], lu = u - g.voff;
            for (int q = 0; q < 3; ++q)
                x[q] = x0[size_t(q) * g.n + lu] + ((m >> q & 1u) ? P.res[size_t(q) * P.n + pu] : 0.0f);
        };
        uint32_t cc[3];
        coords(g, c.local, cc);
        float x[3], d[6], y[3] = {0.0f, 0.0f, 0.0f}, acc[3] = {0.0f, 0.0f, 0.0f};
        xat(v, x);
        for (int q = 0; q < 6; ++q) d[q] = NT ? __builtin_nontemporal_load(A.d32 + q * R + v) : A.d32[q * R + v];
        sym_mul_acc(d, x, y);
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = stride_of(g, a);
            float xn[3], k[6];
            if (cc[a] + 1 < extent_of(g, a)) {
                xat(v + s, xn);
                for (int q = 0; q < 6; ++q) k[q] = A.k32[(size_t(a) * 6 + q) * R + v];
                sym_mul_acc(k, xn, acc);
            }
            if (cc[a] > 0) {
                xat(v - s, xn);
                for (int q = 0; q < 6; ++q) k[q] = A.k32[(size_t(a) * 6 + q) * R + v - s];
                sym_mul_acc(k, xn, acc);
            }
        }
        const uint32_t m = A.mask[v];
        float rr[3], di[6], dz[3] = {0.0f, 0.0f, 0.0f}, z[3];
        for (int q = 0; q < 3; ++q) rr[q] = A.r[q * R + v] - (y[q] - ((m >> q & 1u) ? acc[q] : 0.0f));
        for (int q = 0; q < 6; ++q) di[q] = NT ? __builtin_nontemporal_load(A.dinv32 + q * R + v) : A.dinv32[q * R + v];
        sym_mul_acc(di, rr, dz);
        for (int q = 0; q < 3; ++q) {
            z[q] = __builtin_fmaf(kMgOmega, dz[q], x[q]);
            if constexpr (NT) __builtin_nontemporal_store(z[q], A.z + q * R + v);
            else A.z[q * R + v] = z[q];
        }
        const float r0 = A.r[v], r1 = A.r[R + v], r2 = A.r[2 * R + v];
        part[0] = __builtin_fma(double(r2), double(z[2]), __builtin_fma(double(r1), double(z[1]), double(r0) * double(z[0])));
    }
    const int out[1] = {kRz};
    tile_store<1>(part, A, c.t, out, lds);
}

__global__ void k_zero_word(uint32_t *w) { *w = 0u; }

// Host side: the hierarchy of a one-scene arena (device buffers in one
// allocation, freed with the arena).
struct MgHost {
    std::vector<MgLevel> lv;  // levels 1..L
    float *fres = nullptr;    // the fine level's residual (3n)
    float *fx0 = nullptr;     // the fine level's pre-smoothed iterate (3n)
    double *M = nullptr;  // the coarsest level, dense; after set-up its inverse
    uint32_t *err = nullptr;
    uint32_t n3 = 0;
    void *alloc = nullptr;
    uint64_t built = ~0ull;   // assembly generation the hierarchy was built from
};

}  // namespace
}  // namespace vx

struct vx_batch {
    vx::DevArena A{};
    uint32_t cap_tiles = 0, chunk = 1, max_slots = 0;
    void *alloc = nullptr;
    size_t alloc_bytes = 0;
    std::vector<vx::SlotHost> slots;
    std::vector<vx::SlotGeo> tiles;    // host copy of the per-tile table
    std::vector<vx::SlotState> known;  // last state read back per slot
    std::vector<uint32_t> inflight;    // bits of each slot in the tick in flight
    vx::SlotGeo *d_tiles = nullptr, *d_geo = nullptr;
    uint32_t *d_bits = nullptr;
    vx::SlotParams *d_prm = nullptr;
    vx::SlotState *d_st = nullptr;
    // Two ticks alternate: fl[cur] is filled next, fl[cur ^ 1] may be in flight.
    vx::Flight fl[2];
    int cur = 0;
    // Inner chunk graphs by (precision, launch grid in tiles).
    std::vector<std::pair<uint64_t, hipGraphExec_t>> g_inner;
    bool single = false;  // one-scene arena of a vx_scene (never re-admits)
    uint32_t hw_tiles = 0;  // one past the highest tile in use
    uint32_t iters_per_tick = vx::kInnerChunk;
    // Launch shape of a one-scene arena's inner kernels (vx_scene_set_shape).
    uint32_t zcols = vx::kZColsDefault;
    uint32_t deal = VX_DEAL_AUTO;
    uint32_t update_threads = vx::kBlock;  // threads per tile of the block-Jacobi update
    // Bits and launch grid of the last spread_bits enqueued; a tick with the
    // same bits and grid and an unchanged tile table skips it.
    std::vector<uint32_t> spread_last;
    uint32_t spread_tiles = 0;
    bool spread_dirty = true;
    vx::ToolHost *tool = nullptr;  // the rigid tool of a one-scene arena
    // Tile list of the last tick that had one (its device copy is d_list).
    uint32_t *d_list = nullptr;
    std::vector<uint32_t> list_last;
    // Multigrid hierarchy of a one-scene arena (built on first use) and the
    // count of assemblies enqueued, which tells whether it is current.
    vx::MgHost *mg = nullptr;
    uint64_t asm_gen = 0;
};

struct vx_scene {
    vx_batch *batch = nullptr;
    uint32_t slot = 0;
    uint32_t active = 0;
};

namespace vx {
namespace {

struct Lock {
    std::lock_guard<std::recursive_mutex> g;
    Lock() : g(context().mutex) {}
};

// Per-dispatch timing of the inner kernels (vx_scene_solve_timed): each
// timed launch carries a start and a stop event that the runtime binds to the
// dispatch's own begin and end timestamps (hipExtLaunchKernelGGL), the
// figures a kernel-dispatch trace reports. Only direct launches are timed.
struct DispatchTimer {
    static constexpr int kKinds = 2;  // 0: pcg_bond_spmv, 1: pcg_update
    std::vector<hipEvent_t> pool;     // start/stop pairs
    std::vector<int> kind;            // kind of each pair in use
    size_t used = 0;
    std::vector<float> us[kKinds];
    int32_t take(int k, hipEvent_t *start, hipEvent_t *stop) {
        if (2 * (used + 1) > pool.size()) {
            for (int i = 0; i < 2; ++i) {
                hipEvent_t e = nullptr;
                VX_HIP(hipEventCreate(&e));
                pool.push_back(e);
            }
        }
        if (kind.size() <= used) kind.resize(used + 1);
        kind[used] = k;
        *start = pool[2 * used];
        *stop = pool[2 * used + 1];
        ++used;
        return VX_OK;
    }
    // After the stream has passed every timed dispatch.
    int32_t harvest() {
        for (size_t i = 0; i < used; ++i) {
            float ms = 0.0f;
            VX_HIP(hipEventElapsedTime(&ms, pool[2 * i], pool[2 * i + 1]));
            us[kind[i]].push_back(1e3f * ms);
        }
        used = 0;
        return VX_OK;
    }
    ~DispatchTimer() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};
DispatchTimer *g_timer = nullptr;  // set for the duration of a timed solve (under the context lock)

// Launch `kernel` on the library stream, with dispatch timestamps of kind
// `k` while a timed solve runs.
template <class K, class... Args>
void launch_kernel(int k, K kernel, const dim3 &grid, const dim3 &blk, Args... args) {
    Context &ctx = context();
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (g_timer && k >= 0 && g_timer->take(k, &t0, &t1) == VX_OK)
        hipExtLaunchKernelGGL(kernel, grid, blk, 0, ctx.stream, t0, t1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, blk, 0, ctx.stream, args...);
}

// The multigrid hierarchy of a one-scene arena: level grids and buffers,
// allocated once (the scene's geometry never changes).
int32_t mg_prepare(vx_batch *b) {
    if (b->mg) return VX_OK;
    if (!b->single) return fail(VX_ERR_INVALID_ARGUMENT, "the multigrid preconditioner needs a one-scene arena");
    const SlotGeo &g = b->slots[0].geo;
    std::vector<MgLevel> lv;
    uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
    do {
        nx = (nx + 1) / 2;
        ny = (ny + 1) / 2;
        nz = (nz + 1) / 2;
        MgLevel L{};
        L.nx = nx;
        L.ny = ny;
        L.nz = nz;
        L.n = nx * ny * nz;
        lv.push_back(L);
    } while (uint64_t(nx) * ny * nz > kMgCoarsest && lv.size() < kMgMaxLevels);
    if (lv.back().n > kMgCoarsest) return fail(VX_ERR_TOO_LARGE, "multigrid: the scene is too elongated to coarsen");
    const uint32_t n3 = 3 * lv.back().n;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t bytes = 2 * up(sizeof(float) * 3 * size_t(g.n)) + up(sizeof(double) * size_t(n3) * n3) + up(sizeof(uint32_t));
    for (const MgLevel &L : lv) bytes += up(sizeof(float) * 48 * size_t(L.n)) + up(L.n) + up(sizeof(double) * 33 * size_t(L.n));
    void *mem = nullptr;
    VX_HIP(hipMalloc(&mem, bytes));
    MgHost *h = new MgHost;
    h->alloc = mem;
    char *q = static_cast<char *>(mem);
    auto take = [&](size_t nbytes) {
        char *r = q;
        q += up(nbytes);
        return r;
    };
    h->fres = reinterpret_cast<float *>(take(sizeof(float) * 3 * size_t(g.n)));
    h->fx0 = reinterpret_cast<float *>(take(sizeof(float) * 3 * size_t(g.n)));
    h->M = reinterpret_cast<double *>(take(sizeof(double) * size_t(n3) * n3));
    h->err = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t)));
    for (MgLevel &L : lv) {
        float *f = reinterpret_cast<float *>(take(sizeof(float) * 48 * size_t(L.n)));
        L.D = f;
        L.C = f + 6 * size_t(L.n);
        L.Dinv = f + 33 * size_t(L.n);
        L.r = f + 39 * size_t(L.n);
        L.x = f + 42 * size_t(L.n);
        L.res = f + 45 * size_t(L.n);
        L.mask = reinterpret_cast<uint8_t *>(take(L.n));
        double *d = reinterpret_cast<double *>(take(sizeof(double) * 33 * size_t(L.n)));
        L.D64 = d;
        L.C64 = d + 6 * size_t(L.n);
    }
    h->lv = lv;
    h->n3 = n3;
    b->mg = h;
    return VX_OK;
}
inline dim3 mg_grid(uint32_t n) { return dim3((n + kMgBlock - 1) / kMgBlock); }

// Build the hierarchy from the scene's current FP64 operator (after an
// assembly), on the library stream.
void mg_setup(vx_batch *b) {
    hipStream_t s = context().stream;
    MgHost &h = *b->mg;
    const SlotGeo &g = b->slots[0].geo;
    hipLaunchKernelGGL(mg_build_fine, mg_grid(h.lv[0].n), dim3(kMgBlock), 0, s, b->A, g, h.lv[0]);
    for (size_t l = 1; l < h.lv.size(); ++l)
        hipLaunchKernelGGL(mg_build_coarse, mg_grid(h.lv[l].n), dim3(kMgBlock), 0, s, h.lv[l - 1], h.lv[l]);
    hipLaunchKernelGGL(mg_dense, mg_grid(h.lv.back().n), dim3(kMgBlock), 0, s, h.lv.back(), h.M, h.n3);
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, h.err);
    const MgLevel &Lc = h.lv.back();
    const uint32_t bw = std::min<uint32_t>(h.n3 - 1, 3 * Lc.nx * Lc.ny + 2);  // the stencil's reach in DOF order
    hipLaunchKernelGGL(mg_cholesky, dim3(1), dim3(kMgDenseBlock), 0, s, h.M, h.n3, bw, h.err);
}

// One V-cycle z = M^-1 r on the fine level and the partials of r.z; every
// kernel returns at once when the scene's inner solve has ended.
template <bool NT>
void mg_cycle(vx_batch *b, const OneScene &one_u, bool big) {
    hipStream_t s = context().stream;
    const MgHost &h = *b->mg;
    const SlotGeo &g = b->slots[0].geo;
    const SlotState *st = b->A.st + g.slot;
    const size_t nl = h.lv.size();
    hipLaunchKernelGGL(mg_fine_down<NT>, mg_grid(g.n), dim3(kMgBlock), 0, s, b->A, g, st, h.fres, h.fx0);
    hipLaunchKernelGGL(mg_restrict_fine, mg_grid(h.lv[0].n), dim3(kMgBlock), 0, s, b->A, g, st, h.fres, h.lv[0]);
    for (size_t l = 0; l + 1 < nl; ++l) {
        hipLaunchKernelGGL(mg_coarse_down, mg_grid(h.lv[l].n), dim3(kMgBlock), 0, s, h.lv[l], st);
        hipLaunchKernelGGL(mg_restrict_coarse, mg_grid(h.lv[l + 1].n), dim3(kMgBlock), 0, s, h.lv[l], h.lv[l + 1], st);
    }
    hipLaunchKernelGGL(mg_coarsest, dim3((h.n3 + kMgDenseBlock / 64 - 1) / (kMgDenseBlock / 64)), dim3(kMgDenseBlock), 0, s,
                       h.lv[nl - 1], h.M, h.n3, st);
    for (size_t l = nl - 1; l-- > 0;)
        hipLaunchKernelGGL(mg_coarse_up, mg_grid(h.lv[l].n), dim3(kMgBlock), 0, s, h.lv[l], h.lv[l + 1], st);
    hipLaunchKernelGGL(mg_fine_up<NT>, dim3(one_u.grid), dim3(kBlock), 0, s, b->A, one_u, h.lv[0], h.fx0);
    if (big) hipLaunchKernelGGL(scene_reduce<float>, dim3(b->max_slots), dim3(kBlock), 0, s, b->A, int(kRz), 1, 0);
}

template <class S, bool ONE, bool RB, bool EARLY, bool NT>
void launch_iteration_t(vx_batch *b, const OneScene &one, const dim3 &grid, const OneScene &one_u, const dim3 &grid_u,
                        int odd, bool big, bool mg) {
    Context &ctx = context();
    const dim3 blk(kBlock), sg(b->max_slots);
    auto reduce = [&](int k0, int nk) {
        if (big) hipLaunchKernelGGL(scene_reduce<S>, sg, blk, 0, ctx.stream, b->A, k0, nk, 0);
    };
    // With multigrid the iteration starts with the cycle: z = M^-1 r and the
    // partials of r.z replace those of the update before (whose z = D^-1 r
    // the cycle overwrites).
    if constexpr (ONE && !RB && std::is_same<S, float>::value) {
        if (mg) mg_cycle<NT>(b, one_u, big);
    }
    if constexpr (RB) hipLaunchKernelGGL((pcg_rb_red<S, ONE, EARLY>), grid_u, blk, 0, ctx.stream, b->A, one_u, odd);
    if constexpr (EARLY && !RB && std::is_same<S, float>::value)
        launch_kernel(0, pcg_bond_spmv_big<S, ONE, NT>, grid, blk, b->A, one, odd);
    else
        launch_kernel(0, pcg_bond_spmv<S, ONE, RB, EARLY>, grid, blk, b->A, one, odd);
    reduce(kPap, 1);
    if constexpr (ONE && !RB) {
        if (b->update_threads == kHalf) {
            launch_kernel(1, pcg_update_pair<S, ONE, EARLY>, grid_u, dim3(kHalf), b->A, one_u, odd);
            reduce(kRz, 2);
            return;
        }
    }
    launch_kernel(1, pcg_update<S, ONE, RB, EARLY, NT>, grid_u, blk, b->A, one_u, odd);
    reduce(kRz, 2);
}
// The inner kernels of one iteration for precision S: with red-black scenes
// in the launch (rb), pcg_rb_red first and the kernels that branch per tile.
// One-scene arenas: the SpMV with `one` / `grid`, the other kernels with the
// round-robin mapping `one_u` / `grid_u`.
template <class S, bool EARLY, bool NT>
void launch_iteration_e(vx_batch *b, bool rb, const OneScene &one, const dim3 &grid, const OneScene &one_u,
                        const dim3 &grid_u, int odd, bool big, bool mg) {
    if (rb && b->single)
        launch_iteration_t<S, true, true, EARLY, NT>(b, one, grid, one_u, grid_u, odd, big, false);
    else if (rb)
        launch_iteration_t<S, false, true, EARLY, NT>(b, one, grid, one_u, grid_u, odd, big, false);
    else if (b->single)
        launch_iteration_t<S, true, false, EARLY, NT>(b, one, grid, one_u, grid_u, odd, big, mg);
    else
        launch_iteration_t<S, false, false, EARLY, NT>(b, one, grid, one_u, grid_u, odd, big, false);
}
template <class S>
void launch_iteration(vx_batch *b, bool rb, const OneScene &one, const dim3 &grid, const OneScene &one_u,
                      const dim3 &grid_u, int odd, bool big, bool nt, bool mg) {
    if (big && nt)
        launch_iteration_e<S, true, true>(b, rb, one, grid, one_u, grid_u, odd, big, mg);
    else if (big)
        launch_iteration_e<S, true, false>(b, rb, one, grid, one_u, grid_u, odd, big, mg);
    else
        launch_iteration_e<S, false, false>(b, rb, one, grid, one_u, grid_u, odd, big, mg);
}

// One inner chunk: the refinement residual of scenes whose inner solve has
// ended (device-decided), then the arena's iterations per tick. Enqueued on
// the library stream (captured into a graph, or launched directly).
// The XCD deal of a one-scene arena's SpMV and of the kernels after it.
// Default: the SpMV in contiguous eighths; the others round robin for scenes
// of more than kBigTiles tiles (their inputs come from HBM anyway) and
// contiguous below, so the data a kernel reads was produced on its own XCD.
bool deal_spmv_rr(const vx_batch *b) { return b->deal == VX_DEAL_ROUND_ROBIN; }
bool deal_update_rr(const vx_batch *b, bool big) {
    return b->deal == VX_DEAL_AUTO ? big : b->deal != VX_DEAL_CONTIGUOUS;
}

void enqueue_inner(vx_batch *b, int f64, bool rb, uint32_t tiles, bool big, bool nt, uint32_t iters, bool mg) {
    Context &ctx = context();
    const dim3 grid(tiles), blk(kBlock), sg(b->max_slots);
    if (f64) {
        if (rb) hipLaunchKernelGGL(refine_red<double>, grid, blk, 0, ctx.stream, b->A, 0);
        hipLaunchKernelGGL(refine_residual<double>, grid, blk, 0, ctx.stream, b->A, 0);
        hipLaunchKernelGGL(refine_decide<double>, sg, blk, 0, ctx.stream, b->A, 0);
        if (big) hipLaunchKernelGGL(scene_reduce<double>, sg, blk, 0, ctx.stream, b->A, int(kRz), 2, 0);
    } else {
        if (rb) hipLaunchKernelGGL(refine_red<float>, grid, blk, 0, ctx.stream, b->A, 0);
        hipLaunchKernelGGL(refine_residual<float>, grid, blk, 0, ctx.stream, b->A, 0);
        hipLaunchKernelGGL(refine_decide<float>, sg, blk, 0, ctx.stream, b->A, 0);
        if (big) hipLaunchKernelGGL(scene_reduce<float>, sg, blk, 0, ctx.stream, b->A, int(kRz), 2, 0);
    }
    // A one-scene arena (vx_scene_*) never changes its scene: its geometry
    // goes into the kernel arguments, with its launch shape.
    OneScene one = b->single ? make_one(b->slots[0].geo, deal_spmv_rr(b), b->zcols) : OneScene{};
    OneScene one_u = b->single ? make_one(b->slots[0].geo, deal_update_rr(b, big), b->zcols) : OneScene{};
    one.rb = one_u.rb = rb ? 1u : 0u;
    const dim3 gi(b->single ? one.grid : tiles), gu(b->single ? one_u.grid : tiles);
    for (uint32_t t = 0; t < iters; ++t) {
        if (f64)
            launch_iteration<double>(b, rb, one, gi, one_u, gu, int(t & 1), big, nt, false);
        else
            launch_iteration<float>(b, rb, one, gi, one_u, gu, int(t & 1), big, nt, mg);
    }
}

int32_t capture_inner(vx_batch *b, int f64, bool rb, uint32_t tiles, bool big, bool nt, uint32_t iters, bool mg,
                      hipGraphExec_t *out) {
    Context &ctx = context();
    hipGraph_t graph;
    VX_HIP(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
    enqueue_inner(b, f64, rb, tiles, big, nt, iters, mg);
    VX_HIP(hipStreamEndCapture(ctx.stream, &graph));
    hipError_t e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return VX_OK;
}

// Workgroups per arena-wide launch: the tiles in use rounded up to a whole
// number of XCD rounds (8 * chunk) times a power of two, so launches skip the
// free end of a large arena and only a few chunk graphs are ever captured.
uint32_t launch_tiles(const vx_batch *b) {
    const uint64_t unit = uint64_t(kXcds) * b->chunk;
    uint64_t g = unit;
    while (g < b->hw_tiles) g *= 2;
    // Stay a whole number of XCD rounds (logical_tile is a bijection on
    // [0, g) only then); tiles past the arena's capacity exit at once.
    return uint32_t(std::min<uint64_t>(g, (b->cap_tiles + unit - 1) / unit * unit));
}

// Whether any admitted scene takes its inner partials through scene_reduce.
bool any_big(const vx_batch *b) {
    for (const auto &h : b->slots)
        if (h.used && h.geo.ntiles > kBigTiles) return true;
    return false;
}

// Whether the FP32 inner kernels of large arenas stream their planes
// (nontemporal loads of the planes a launch reads once, nontemporal stores):
// when the tiles in use hold more inner-plane bytes (about 205 B per row)
// than the 256 MB Infinity Cache, nothing survives in it from one kernel to
// the next, and streaming keeps the SpMV's re-read neighbour planes in the
// L2. Below that, the planes stay in the cache between kernels and the
// ordinary accesses are faster (96x96x64: 58.3 ms block-Jacobi solves
// against 62.4 ms streamed; 192x192x96: 192.8 against 177.3 ms).
constexpr uint32_t kStreamTiles = 2560;  // 1.31 M rows
bool streams(const vx_batch *b) { return b->hw_tiles >= kStreamTiles; }

int32_t inner_graph(vx_batch *b, int f64, bool rb, uint32_t tiles, uint32_t iters, bool mg, hipGraphExec_t *out) {
    const bool big = any_big(b), nt = big && streams(b);
    const uint64_t key = (uint64_t(iters) << 48) | (uint64_t(mg) << 47) | (uint64_t(nt) << 46) | (uint64_t(tiles) << 4) |
                         (uint64_t(b->A.tile_list != nullptr) << 3) | (uint64_t(rb) << 2) | (uint64_t(big) << 1) |
                         uint64_t(f64);
    for (auto &kv : b->g_inner)
        if (kv.first == key) {
            *out = kv.second;
            return VX_OK;
        }
    if (int32_t c = capture_inner(b, f64, rb, tiles, big, nt, iters, mg, out)) return c;
    b->g_inner.emplace_back(key, *out);
    return VX_OK;
}

// Inner iterations of the next chunk of a one-scene arena. A chunk's head
// replaces the residual when the segment before it has ended, so the
// iterations of a chunk after its segment's end do nothing (their launches
// still load their rows). When the last state read back shows the inner
// recurrence at the solve's tolerance, the solve most likely ends at the next
// head: the chunk gets a short body. Otherwise, from the last two residuals
// read back in the segment, the segment's end is extrapolated at the
// per-iteration rate of its last stretch (CG speeds up within a segment, so
// the latest rate is the closest guess). When the end falls inside the next
// chunk (after the chunk in flight), the chunk is cut to end two iterations
// after it. When it falls inside the chunk in flight, that chunk either was
// cut for it already (the next chunk then starts the new segment: full
// length) or is a full chunk, and the next chunk is kept short to catch an end
// that comes later than guessed. A wrong guess costs a chunk boundary (too
// short) or part of a chunk spent after the end (too long), never a result
// bit: the iterations are those of the same device-decided recurrence either
// way.
uint32_t chunk_iters(SlotHost &h, const SlotState &st, uint32_t full, uint32_t in_flight) {
    constexpr uint32_t kCatch = 8;  // the short chunk after a full one that should hold the end
    if (st.done != kRun) {
        h.seen_seg = kNoSeg;
        const double tol = h.params.solve.tolerance;
        if (st.done == kConv && st.rr_last <= tol * tol * st.bb) return std::min(4u, full);
        return full;
    }
    const uint32_t k = st.k_update;
    const double rr = st.rr_last;
    if (h.seen_seg != st.seg || k <= h.seen_k || !(rr > 0.0) || !(h.seen_rr > 0.0)) {
        h.seen_seg = st.seg;
        h.seen_k = k;
        h.seen_rr = rr;
        return full;
    }
    // per iteration, < 0 when converging: the last stretch's, or the
    // segment's so far when the last stretch did not converge
    const double rate_last = std::log(rr / h.seen_rr) / double(k - h.seen_k);
    const double rate_seg = k > st.seg && st.rr0 > 0.0 ? std::log(rr / st.rr0) / double(k - st.seg) : rate_last;
    const double rate = rate_last < 0.0 ? rate_last : rate_seg;
    h.seen_k = k;
    h.seen_rr = rr;
    const double it = h.params.solve.inner_tolerance, tol = h.params.solve.tolerance;
    const double target = std::max(it * it * st.rr0, tol * tol * st.bb);
    if (!(rate < 0.0) || !(target > 0.0) || !(rr > target)) return full;
    const double left = std::log(target / rr) / rate - double(in_flight);  // after the chunk in flight
    if (!(left < double(full))) return full;
    if (!(left > 0.0)) return in_flight > 0 && in_flight < full ? full : std::min(kCatch, full);
    const uint32_t c = (uint32_t(left + 2.0) + 3u) & ~3u;  // two past the guess, a multiple of 4
    return std::max(4u, std::min(c, full));
}

// Enqueue one tick for the bits in f.bits: the pre-inner kernels, one inner
// chunk per precision in use, the post-inner kernels, then the read-back of
// all scene states into f.st. Does not wait.
int32_t launch_tick(vx_batch *b, Flight &f) {
    Context &ctx = context();
    hipStream_t s = ctx.stream;
    uint32_t any = 0;
    bool inner[2] = {false, false}, resid0[2] = {false, false}, rb[2] = {false, false}, rb0[2] = {false, false};
    bool mg = false;  // the one scene solves with the multigrid cycle (FP32 inner only)
    for (uint32_t i = 0; i < b->max_slots; ++i) {
        if (!b->slots[i].used) f.bits[i] = 0;
        const uint32_t bi = f.bits[i];
        any |= bi;
        if ((bi & B_MG) && (bi & (B_INNER | B_RESID0)) && !(bi & B_F64)) mg = true;
        const int prec = (bi & B_F64) ? 1 : 0;
        const bool red_black = bi & B_RB;
        if (bi & B_INNER) {
            inner[prec] = true;
            rb[prec] = rb[prec] || red_black;
        }
        if (bi & B_RESID0) {
            resid0[prec] = true;
            rb0[prec] = rb0[prec] || red_black;
        }
    }
    // Tiles in use (rounded up), the physical range the tile bits cover; the
    // launches cover it too unless the tick's scenes hold at most half of it,
    // in which case they take the list of those scenes' tiles.
    const uint32_t span = launch_tiles(b);
    uint32_t tiles = span;
    b->A.tile_list = nullptr;
    if (!b->single) {
        uint64_t act = 0;
        std::vector<std::pair<uint32_t, uint32_t>> runs;  // (first tile, tiles) of the tick's scenes
        for (uint32_t i = 0; i < b->max_slots; ++i)
            if (f.bits[i] && b->slots[i].used) {
                runs.emplace_back(b->slots[i].geo.tile0, b->slots[i].geo.ntiles);
                act += b->slots[i].geo.ntiles;
            }
        const uint64_t unit = uint64_t(kXcds) * b->chunk;
        uint64_t g = unit;
        while (g < act) g *= 2;
        if (2 * g <= span) {
            std::sort(runs.begin(), runs.end());
            uint32_t k = 0;
            for (const auto &r : runs)
                for (uint32_t t = 0; t < r.second; ++t) f.list[k++] = r.first + t;
            while (k < g) f.list[k++] = kNoTile;
            if (b->list_last.size() != g || !std::equal(f.list, f.list + g, b->list_last.begin())) {
                VX_HIP(hipMemcpyAsync(b->d_list, f.list, sizeof(uint32_t) * g, hipMemcpyHostToDevice, s));
                b->list_last.assign(f.list, f.list + g);
            }
            b->A.tile_list = b->d_list;
            tiles = uint32_t(g);
        }
    }
    // every kernel of this tick takes the arena with or without the list; the
    // arena is reset to no list once they are enqueued
    struct ListReset {
        DevArena &A;
        ~ListReset() { A.tile_list = nullptr; }
    } list_reset{b->A};
    const dim3 grid(tiles), blk(kBlock), sg(b->max_slots);
    if (b->spread_dirty || span != b->spread_tiles || !std::equal(f.bits, f.bits + b->max_slots, b->spread_last.begin())) {
        BitsArg arg{};
        if (b->max_slots <= kArgSlots) {
            arg.n = b->max_slots;
            std::copy(f.bits, f.bits + b->max_slots, arg.bits);
        } else {
            VX_HIP(hipMemcpyAsync(b->d_bits, f.bits, sizeof(uint32_t) * b->max_slots, hipMemcpyHostToDevice, s));
        }
        const uint32_t nb = std::max((span + kBlock - 1) / kBlock, (arg.n + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(spread_bits, dim3(nb), blk, 0, s, b->A, arg);
        std::copy(f.bits, f.bits + b->max_slots, b->spread_last.begin());
        b->spread_tiles = span;
        b->spread_dirty = false;
    }
    if (any & B_PREDICT) hipLaunchKernelGGL(substep_predict, grid, blk, 0, s, b->A);
    // The rigid tool of a one-scene arena: one advance over dt (0: the
    // zero-time pass).
    auto tool_pass = [&](bool timed) {
        const SlotHost &sh = b->slots[0];
        const double dt = timed ? sh.params.fixed_dt / double(sh.params.substeps) : 0.0;
        const ToolBufs &tb = b->tool->bufs;
        hipLaunchKernelGGL(tool_prepare, dim3(1), dim3(1), 0, s, tb, dt, sh.params.gravity[0], sh.params.gravity[1],
                           sh.params.gravity[2], uint32_t(timed));
        hipLaunchKernelGGL(tool_filter_count, grid, blk, 0, s, b->A, tb);
        hipLaunchKernelGGL(tool_filter_scan, dim3(1), dim3(kToolBlock), 0, s, tb, sh.geo.ntiles);
        hipLaunchKernelGGL(tool_filter_write, grid, blk, 0, s, b->A, tb);
        hipLaunchKernelGGL(tool_advance, dim3(1), dim3(kToolBlock), 0, s, b->A, tb, dt, uint32_t(timed), sh.geo);
    };
    // The tool advances over the substep after the prediction
    // (src/solver.rs:530-543).
    if ((any & B_TOOL) && (any & B_PREDICT) && b->tool) tool_pass(true);
    if (any & B_FIT) hipLaunchKernelGGL(fit_rotations, grid, blk, 0, s, b->A, uint32_t(B_FIT), 0);
    if (any & B_ASSEMBLE) {
        hipLaunchKernelGGL(assemble_bonds, grid, blk, 0, s, b->A);
        hipLaunchKernelGGL(assemble_rows, grid, blk, 0, s, b->A);
        b->asm_gen += 1;
    }
    if (mg) {
        // the hierarchy follows the latest assembly
        if (int32_t c = mg_prepare(b)) return c;
        if (b->mg->built != b->asm_gen) {
            mg_setup(b);
            b->mg->built = b->asm_gen;
        }
    }
    const bool big = any_big(b);
    if (resid0[0]) {
        if (rb0[0]) hipLaunchKernelGGL(refine_red<float>, grid, blk, 0, s, b->A, 1);
        hipLaunchKernelGGL(refine_residual<float>, grid, blk, 0, s, b->A, 1);
        hipLaunchKernelGGL(refine_decide<float>, sg, blk, 0, s, b->A, 1);
        if (big) hipLaunchKernelGGL(scene_reduce<float>, sg, blk, 0, s, b->A, int(kRz), 2, 1);
    }
    if (resid0[1]) {
        if (rb0[1]) hipLaunchKernelGGL(refine_red<double>, grid, blk, 0, s, b->A, 1);
        hipLaunchKernelGGL(refine_residual<double>, grid, blk, 0, s, b->A, 1);
        hipLaunchKernelGGL(refine_decide<double>, sg, blk, 0, s, b->A, 1);
        if (big) hipLaunchKernelGGL(scene_reduce<double>, sg, blk, 0, s, b->A, int(kRz), 2, 1);
    }
    VX_HIP(hipGetLastError());
    f.iters = (inner[0] || inner[1]) ? b->iters_per_tick : 0u;
    if (b->single && (inner[0] || inner[1])) {
        const Flight &in_flight = b->fl[b->cur ^ 1];
        // with multigrid a segment takes about a dozen iterations: shorter chunks
        const uint32_t full = mg ? std::min(b->iters_per_tick, kMgChunk) : b->iters_per_tick;
        f.iters = chunk_iters(b->slots[0], b->known[0], full, in_flight.active ? in_flight.iters : 0u);
    }
    for (int prec = 0; prec < 2; ++prec) {
        if (!inner[prec]) continue;
        const bool mg_prec = mg && prec == 0;
        if (ctx.direct_launch || g_timer) {
            enqueue_inner(b, prec, rb[prec], tiles, big, big && streams(b), f.iters, mg_prec);
            VX_HIP(hipGetLastError());
            continue;
        }
        hipGraphExec_t g = nullptr;
        if (int32_t c = inner_graph(b, prec, rb[prec], tiles, f.iters, mg_prec, &g)) return c;
        VX_HIP(hipGraphLaunch(g, s));
    }
    if (any & (B_APPLY | B_WARM_SAVE)) hipLaunchKernelGGL(apply_correction, grid, blk, 0, s, b->A);
    if (any & B_FITB) {
        hipLaunchKernelGGL(fit_rotations, grid, blk, 0, s, b->A, uint32_t(B_FITB), 1);
        hipLaunchKernelGGL(reduce_angle, sg, blk, 0, s, b->A);
    }
    if (any & B_FINISH) hipLaunchKernelGGL(bond_forces, grid, blk, 0, s, b->A);
    f.frac_timed = (any & B_FRACTURE) != 0;
    if (f.frac_timed) {
        VX_HIP(hipEventRecord(f.frac0, s));
        hipLaunchKernelGGL(fracture_bonds, grid, blk, 0, s, b->A);
        hipLaunchKernelGGL(reduce_fracture, sg, blk, 0, s, b->A);
        VX_HIP(hipEventRecord(f.frac1, s));
    }
    if (any & B_FINISH) hipLaunchKernelGGL(finish_velocities, grid, blk, 0, s, b->A);
    // Its zero-time pass on the substep's final state (src/solver.rs:701-706).
    if ((any & B_TOOL) && (any & B_FINISH) && b->tool) tool_pass(false);
    VX_HIP(hipGetLastError());
    hipLaunchKernelGGL(publish_state, dim3(1), blk, 0, s, b->d_st, f.st_dev, b->max_slots);
    VX_HIP(hipGetLastError());
    VX_HIP(hipEventRecord(f.done, s));
    f.active = true;
    return VX_OK;
}

// Bits of every solve tick: the inner precision and the preconditioner.
uint32_t solve_bits(const vx_solve_params &s) {
    return (s.mode == VX_SOLVE_F64 ? uint32_t(B_F64) : 0u) | (s.preconditioner == VX_PRECOND_RED_BLACK ? uint32_t(B_RB) : 0u) |
           (s.preconditioner == VX_PRECOND_MULTIGRID ? uint32_t(B_MG) : 0u);
}
uint32_t prec_bits(const vx_step_params &p) { return solve_bits(p.solve); }
// The end of a substep: the fracture pass (when on) and the velocities.
// The last tick of a substep: velocities, fracture if enabled, the tool's
// zero-time pass if the scene has one.
uint32_t finish_bits(const SlotHost &h) {
    return uint32_t(B_FINISH) | (h.params.fracture.enabled ? uint32_t(B_FRACTURE) : 0u) | (h.has_tool ? uint32_t(B_TOOL) : 0u);
}

// Bits of an automatically driven scene for its next tick, from its latest
// read-back state.
uint32_t next_bits(const SlotHost &h, const SlotState &st) {
    const uint32_t prec = prec_bits(h.params);
    switch (h.phase) {
    case P_START:
    case P_FINISHED: {
        uint32_t b = B_PREDICT | B_FIT | B_ASSEMBLE | B_RESID0 | B_INNER | prec;
        if (h.has_tool) b |= B_TOOL;
        if (h.params.solve.warm_start && h.warm_ok) b |= B_WARM_USE;
        return b;
    }
    case P_SOLVE:
        if (st.outer_done) {
            uint32_t b = B_APPLY | prec;
            if (h.pass == 0 && h.params.solve.warm_start) b |= B_WARM_SAVE;
            // The reference refits only when another pass may follow
            // (src/solver.rs:561-564).
            b |= (h.pass + 1 < h.params.corotation_iterations) ? uint32_t(B_FITB) : finish_bits(h);
            return b;
        }
        return B_INNER | prec;
    case P_APPLIED:
        if (h.finish_next) return finish_bits(h);
        return B_ASSEMBLE | B_RESID0 | B_INNER | prec;
    default:
        return 0;
    }
}

void finish_substep(SlotHost &h) {
    h.pass = 0;
    h.finish_next = false;
    h.substep += 1;
    if (h.substep == h.params.substeps) {
        h.substep = 0;
        h.steps_done += 1;
    }
    h.phase = h.steps_done >= h.steps ? P_DONE : P_FINISHED;
}

// Advance an automatically driven scene after a tick that ran `bits`.
void advance(SlotHost &h, const SlotState &st, uint32_t bits) {
    if (bits & (B_RESID0 | B_INNER)) h.last_rel = st.rel;
    if (bits & B_FRACTURE) {
        h.frac.plastic_heat += st.frac[0];
        h.frac.stored_plastic += st.frac[1];
        h.frac.surface_work += st.frac[2];
        h.frac.damage_heat += st.frac[3];
        h.frac.released += st.frac[4];
        h.frac.bonds_broken += uint64_t(st.frac[5]);
    }
    if (st.nonfinite || st.done == kNonFin) {
        h.phase = P_FAILED;
        h.status = VX_ERR_NONFINITE_RESIDUAL;
        return;
    }
    if (bits & B_APPLY) {
        h.solves += 1;
        h.inner += st.inner_total;
        h.sweeps += st.sweeps;
        if (h.pass == 0 && h.params.solve.warm_start) h.warm_ok = true;
        if (bits & B_FINISH) {
            finish_substep(h);
            return;
        }
        // Refitted after apply: stop the passes early when the rotations and
        // the solve have both settled (src/solver.rs:565-569).
        h.last_angle = st.angle;
        h.pass += 1;
        h.phase = P_APPLIED;
        h.finish_next = st.angle <= h.params.corotation_tolerance && st.rel <= h.params.solve.tolerance;
        return;
    }
    if (bits & B_FINISH) {
        finish_substep(h);
        return;
    }
    if (bits & (B_PREDICT | B_ASSEMBLE)) h.phase = P_SOLVE;
    if (h.phase == P_SOLVE && st.outer_done && !(st.rel <= h.params.solve.tolerance)) {
        h.phase = P_FAILED;
        h.status = VX_ERR_NOT_CONVERGED;
    }
}

// Wait for a launched tick and take its states in: the scenes it ran adopt
// their new state and automatic scenes advance their phase.
int32_t retire(vx_batch *b, Flight &f) {
    if (!f.active) return VX_OK;
    f.active = false;
    VX_HIP(hipEventSynchronize(f.done));
    float frac_ms = 0.0f;
    if (f.frac_timed) VX_HIP(hipEventElapsedTime(&frac_ms, f.frac0, f.frac1));
    for (uint32_t i = 0; i < b->max_slots; ++i) {
        if (!f.bits[i] || !b->slots[i].used) continue;
        // The fracture launches are shared by every scene in the tick.
        if (f.bits[i] & B_FRACTURE) b->slots[i].frac.seconds += 1e-3 * double(frac_ms);
        b->known[i] = f.st[i];
        if (!b->slots[i].manual) advance(b->slots[i], b->known[i], f.bits[i]);
    }
    return VX_OK;
}

// Retire the tick in flight, if any; afterwards nothing of this arena runs.
int32_t drain(vx_batch *b) {
    Flight &f = b->fl[b->cur ^ 1];
    if (int32_t c = retire(b, f)) return c;
    std::fill(b->inflight.begin(), b->inflight.end(), 0u);
    return VX_OK;
}

// A synchronous tick with fixed bits for one slot.
int32_t manual_tick(vx_batch *b, uint32_t slot, uint32_t bits) {
    if (int32_t c = drain(b)) return c;
    Flight &f = b->fl[b->cur];
    std::fill(f.bits, f.bits + b->max_slots, 0u);
    f.bits[slot] = bits;
    if (int32_t c = launch_tick(b, f)) return c;
    return retire(b, f);
}

int32_t create_batch(uint64_t capacity_rows, uint32_t max_scenes, uint32_t chunk, vx_batch **out) {
    if (!out || capacity_rows == 0 || max_scenes == 0) return fail(VX_ERR_INVALID_ARGUMENT, "invalid batch size");
    *out = nullptr;
    const uint64_t tiles = tiles_for(capacity_rows);
    // Default interleave: a one-scene arena deals each XCD one contiguous
    // eighth of the scene; a multi-scene arena deals 16-tile chunks so its
    // launches can stop at the tiles in use.
    if (chunk == 0) {
        const uint64_t eighth = std::max<uint64_t>(1, (tiles + kXcds - 1) / kXcds);
        chunk = uint32_t(max_scenes == 1 ? eighth : std::min<uint64_t>(eighth, 16));
    }
    const uint64_t unit = uint64_t(kXcds) * chunk;
    const uint64_t cap = (tiles + unit - 1) / unit * unit;
    if (cap * kBlock >= (1ull << 29))
        return fail(VX_ERR_TOO_LARGE, "arena of %llu rows exceeds the index range", (unsigned long long)(cap * kBlock));
    if (int32_t c = ensure_init()) return c;
    Lock lock;
    vx_batch *b = new vx_batch();
    b->cap_tiles = uint32_t(cap);
    b->chunk = chunk;
    b->max_slots = max_scenes;
    b->slots.assign(max_scenes, SlotHost{});
    SlotGeo free_tile{};
    free_tile.slot = kFreeTile;
    b->tiles.assign(b->cap_tiles, free_tile);
    b->known.assign(max_scenes, SlotState{});
    b->inflight.assign(max_scenes, 0u);
    b->spread_last.assign(max_scenes, 0u);
    DevArena &A = b->A;
    const size_t R = size_t(b->cap_tiles) * kBlock;
    A.R = uint32_t(R);
    A.cap_tiles = b->cap_tiles;
    A.chunk = chunk;
    A.max_slots = max_scenes;
    struct Item {
        void **ptr;
        size_t bytes;
    };
    std::vector<Item> items = {
        {(void **)&A.pos, 24 * R}, {(void **)&A.pred, 24 * R}, {(void **)&A.prev, 24 * R}, {(void **)&A.vel, 24 * R},
        {(void **)&A.drift, 24 * R}, {(void **)&A.rot, 32 * R}, {(void **)&A.xa, 24 * R}, {(void **)&A.xb, 24 * R},
        {(void **)&A.xw, 24 * R}, {(void **)&A.targets, 24 * R}, {(void **)&A.mass, 4 * R}, {(void **)&A.young, 4 * R},
        {(void **)&A.young_perp, 4 * R}, {(void **)&A.shear, 4 * R}, {(void **)&A.solid, 4 * R}, {(void **)&A.fiber, 12 * R},
        {(void **)&A.aniso, 4 * R}, {(void **)&A.dtemp, 4 * R}, {(void **)&A.a_par, 4 * R}, {(void **)&A.a_perp, 4 * R},
        {(void **)&A.loads, 12 * R}, {(void **)&A.kinematic, R}, {(void **)&A.bonds, R}, {(void **)&A.tmask, R},
        {(void **)&A.mask, R}, {(void **)&A.damage, 12 * R}, {(void **)&A.kscale, 12 * R}, {(void **)&A.offset, 36 * R},
        {(void **)&A.plastic, 36 * R}, {(void **)&A.k64, 144 * R}, {(void **)&A.force, 72 * R}, {(void **)&A.lambda, 72 * R}, {(void **)&A.d64, 48 * R},
        {(void **)&A.dinv64, 48 * R}, {(void **)&A.rhs, 24 * R}, {(void **)&A.k32, 72 * R}, {(void **)&A.d32, 24 * R},
        {(void **)&A.dinv32, 24 * R}, {(void **)&A.d, 12 * R}, {(void **)&A.r, 12 * R}, {(void **)&A.z, 12 * R},
        {(void **)&A.p0, 12 * R}, {(void **)&A.p1, 12 * R}, {(void **)&A.ap, 12 * R}, {(void **)&A.dd, 24 * R},
        {(void **)&A.rd, 24 * R}, {(void **)&A.zd, 24 * R}, {(void **)&A.p0d, 24 * R}, {(void **)&A.p1d, 24 * R},
        {(void **)&A.apd, 24 * R}, {(void **)&A.strength, 4 * kStrength * R}, {(void **)&A.hist, 12 * kHistory * R},
        {(void **)&A.rot_asm, 32 * R}, {(void **)&A.partials, 8 * kSlots * size_t(b->cap_tiles)},
        {(void **)&b->d_tiles, sizeof(SlotGeo) * b->cap_tiles}, {(void **)&A.tile_bits, 4 * size_t(b->cap_tiles)},
        {(void **)&b->d_bits, 4 * size_t(max_scenes)}, {(void **)&b->d_list, 4 * size_t(b->cap_tiles)},
        {(void **)&b->d_geo, sizeof(SlotGeo) * max_scenes}, {(void **)&b->d_prm, sizeof(SlotParams) * max_scenes},
        {(void **)&b->d_st, sizeof(SlotState) * max_scenes},
    };
    size_t total = 0;
    for (auto &it : items) total += (it.bytes + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&b->alloc, total);
    if (e != hipSuccess) {
        delete b;
        return fail(e == hipErrorOutOfMemory ? VX_ERR_TOO_LARGE : VX_ERR_DEVICE, "arena allocation of %zu bytes failed: %s", total,
                    hipGetErrorString(e));
    }
    b->alloc_bytes = total;
    size_t off = 0;
    for (auto &it : items) {
        *it.ptr = static_cast<char *>(b->alloc) + off;
        off += (it.bytes + 255) & ~size_t(255);
    }
    A.tiles = b->d_tiles;
    A.geo = b->d_geo;
    A.prm = b->d_prm;
    A.bits = b->d_bits;
    A.st = b->d_st;
    Context &ctx = context();
    // Every slot starts idle: no inner work (done = converged) until its
    // first residual.
    std::vector<SlotState> st0(max_scenes);
    for (auto &s : st0) s.done = kConv;
    bool ok = hipMemcpy(b->d_tiles, b->tiles.data(), sizeof(SlotGeo) * b->cap_tiles, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b->d_st, st0.data(), sizeof(SlotState) * max_scenes, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemsetAsync(b->d_geo, 0, sizeof(SlotGeo) * max_scenes, ctx.stream) == hipSuccess &&
              hipMemsetAsync(b->d_bits, 0, 4 * size_t(max_scenes), ctx.stream) == hipSuccess &&
              hipMemsetAsync(A.tile_bits, 0, 4 * size_t(b->cap_tiles), ctx.stream) == hipSuccess &&
              hipMemsetAsync(A.partials, 0, 8 * kSlots * size_t(b->cap_tiles), ctx.stream) == hipSuccess &&
              hipStreamSynchronize(ctx.stream) == hipSuccess;
    for (auto &f : b->fl) {
        ok = ok && hipHostMalloc((void **)&f.bits, sizeof(uint32_t) * max_scenes, hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void **)&f.list, sizeof(uint32_t) * b->cap_tiles, hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void **)&f.st, sizeof(SlotState) * max_scenes, hipHostMallocCoherent) == hipSuccess &&
             hipHostGetDevicePointer((void **)&f.st_dev, f.st, 0) == hipSuccess &&
             hipEventCreateWithFlags(&f.done, hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&f.frac0) == hipSuccess && hipEventCreate(&f.frac1) == hipSuccess;
        if (f.bits) std::memset(f.bits, 0, sizeof(uint32_t) * max_scenes);
    }
    b->known = st0;
    if (!ok) {
        for (auto &f : b->fl) {
            if (f.bits) (void)hipHostFree(f.bits);
            if (f.list) (void)hipHostFree(f.list);
            if (f.st) (void)hipHostFree(f.st);
            if (f.done) (void)hipEventDestroy(f.done);
            if (f.frac0) (void)hipEventDestroy(f.frac0);
            if (f.frac1) (void)hipEventDestroy(f.frac1);
        }
        (void)hipFree(b->alloc);
        delete b;
        return fail(VX_ERR_DEVICE, "arena set-up failed");
    }
    *out = b;
    return VX_OK;
}

// Per REV: targeted axes (tmask) and solved axes (mask: not kinematic, not
// targeted); active = REVs with a solved axis (src/solver.rs:1554-1565).
// Targets must be finite where they are not NaN (src/solver.rs:455-459).
int32_t free_masks(const uint8_t *kinematic, const float *targets, uint32_t n, std::vector<uint8_t> &tmask,
                   std::vector<uint8_t> &mask, uint32_t *active) {
    tmask.assign(n, 0);
    mask.assign(n, 0);
    uint32_t act = 0;
    for (uint32_t v = 0; v < n; ++v) {
        uint8_t m = 0;
        for (int a = 0; a < 3; ++a) {
            const float t = targets ? targets[3 * size_t(v) + a] : NAN;
            const bool targeted = !std::isnan(t);
            if (targeted && !std::isfinite(t)) return fail(VX_ERR_INVALID_ARGUMENT, "target of REV %u axis %d is not finite", v, a);
            if (targeted) tmask[v] |= uint8_t(1u << a);
            if (!kinematic[v] && !targeted) m |= uint8_t(1u << a);
        }
        mask[v] = m;
        act += m ? 1 : 0;
    }
    *active = act;
    return VX_OK;
}

// Validate a box description; fills per-REV bond / target / free masks.
int32_t check_desc(const vx_scene_desc *d, std::vector<uint8_t> &bonds, std::vector<uint8_t> &tmask,
                   std::vector<uint8_t> &mask, uint32_t *active) {
    if (!d) return fail(VX_ERR_INVALID_ARGUMENT, "null scene description");
    const uint64_t n64 = uint64_t(d->nx) * d->ny * d->nz;
    if (n64 == 0) return fail(VX_ERR_INVALID_ARGUMENT, "empty scene");
    if (n64 >= (1ull << 31)) return fail(VX_ERR_TOO_LARGE, "scene of %llu REVs exceeds the index range", (unsigned long long)n64);
    if (!d->positions || !d->rotations || !d->velocities || !d->mass || !d->kinematic || !d->young || !d->young_perp ||
        !d->shear || !d->solid || !d->fibers || !d->anisotropy || !d->delta_temp || !d->alpha_par || !d->alpha_perp ||
        !d->neighbor)
        return fail(VX_ERR_INVALID_ARGUMENT, "missing scene array");
    const uint32_t n = uint32_t(n64), nx = d->nx, ny = d->ny;
    bonds.assign(n, 0);
    for (uint32_t a = 0; a < 3; ++a) {
        const uint32_t s = a == 0 ? 1 : (a == 1 ? nx : nx * ny);
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t j = d->neighbor[size_t(a) * n + v];
            if (j == VX_EMPTY) continue;
            const uint32_t c = a == 0 ? v % nx : (a == 1 ? (v / nx) % ny : v / (nx * ny));
            const uint32_t e = a == 0 ? nx : (a == 1 ? ny : d->nz);
            if (c + 1 >= e || j != v + s) return fail(VX_ERR_INVALID_ARGUMENT, "bond %u of axis %u is not a lattice bond", v, a);
            bonds[v] |= uint8_t(1u << a);
        }
    }
    return free_masks(d->kinematic, d->targets, n, tmask, mask, active);
}

// First fit of `need` contiguous free tiles; cap_tiles when none.
uint32_t find_tiles(const vx_batch *b, uint32_t need) {
    uint32_t run = 0;
    for (uint32_t t = 0; t < b->cap_tiles; ++t) {
        run = b->tiles[t].slot == kFreeTile ? run + 1 : 0;
        if (run == need) return t + 1 - need;
    }
    return b->cap_tiles;
}

int32_t put(void *dst, const void *src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? VX_OK : fail(VX_ERR_DEVICE, "scene upload failed: %s", hipGetErrorString(e));
}

// Set the per-tile table entries of tiles [t0, t0 + count) to `g` (or free).
int32_t set_tiles(vx_batch *b, uint32_t t0, uint32_t count, const SlotGeo *g) {
    SlotGeo free_tile{};
    free_tile.slot = kFreeTile;
    for (uint32_t t = t0; t < t0 + count; ++t) b->tiles[t] = g ? *g : free_tile;
    uint32_t hw = b->hw_tiles;
    if (g)
        hw = std::max(hw, t0 + count);
    else
        while (hw > 0 && b->tiles[hw - 1].slot == kFreeTile) --hw;
    b->hw_tiles = hw;
    b->spread_dirty = true;
    return put(b->d_tiles + t0, b->tiles.data() + t0, sizeof(SlotGeo) * count);
}

int32_t upload_scene(vx_batch *b, const vx_scene_desc *d, const SlotGeo &g, const std::vector<uint8_t> &bonds,
                     const std::vector<uint8_t> &tmask, const std::vector<uint8_t> &mask) {
    DevArena &A = b->A;
    const size_t R = A.R, n = g.n, voff = g.voff;
    std::vector<double> buf(n);
    std::vector<float> fbuf(n);
    // f32 AoS (comps per REV) -> f64 planes
    auto up_d = [&](double *dst, const float *src, int comps, bool nan_to_zero) -> int32_t {
        for (int q = 0; q < comps; ++q) {
            for (size_t v = 0; v < n; ++v) {
                const double x = src ? double(src[comps * v + q]) : 0.0;
                buf[v] = (nan_to_zero && std::isnan(x)) ? 0.0 : x;
            }
            if (int32_t c = put(dst + q * R + voff, buf.data(), 8 * n)) return c;
        }
        return VX_OK;
    };
    auto up_f = [&](float *dst, const float *src, int comps, float dflt) -> int32_t {
        for (int q = 0; q < comps; ++q) {
            for (size_t v = 0; v < n; ++v) fbuf[v] = src ? src[comps * v + q] : dflt;
            if (int32_t c = put(dst + q * R + voff, fbuf.data(), 4 * n)) return c;
        }
        return VX_OK;
    };
    // per-bond [axis][voxel] scalars and [axis][voxel][3] vectors
    auto up_bond = [&](float *dst, const float *src, float dflt) -> int32_t {
        for (int a = 0; a < 3; ++a) {
            for (size_t v = 0; v < n; ++v) fbuf[v] = src ? src[a * n + v] : dflt;
            if (int32_t c = put(dst + a * R + voff, fbuf.data(), 4 * n)) return c;
        }
        return VX_OK;
    };
    auto up_bond_vec = [&](float *dst, const float *src) -> int32_t {
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < 3; ++q) {
                for (size_t v = 0; v < n; ++v) fbuf[v] = src ? src[(a * n + v) * 3 + q] : 0.0f;
                if (int32_t c = put(dst + (a * 3 + q) * R + voff, fbuf.data(), 4 * n)) return c;
            }
        return VX_OK;
    };
    int32_t rc;
#define VX_UP(x)                   \
    do {                           \
        if ((rc = (x))) return rc; \
    } while (0)
    VX_UP(up_d(A.pos, d->positions, 3, false));
    VX_UP(up_d(A.pred, d->positions, 3, false));
    VX_UP(up_d(A.prev, d->positions, 3, false));
    VX_UP(up_d(A.vel, d->velocities, 3, false));
    VX_UP(up_d(A.drift, nullptr, 3, false));
    VX_UP(up_d(A.rot, d->rotations, 4, false));
    VX_UP(up_d(A.targets, d->targets, 3, true));
    VX_UP(up_d(A.xa, nullptr, 3, false));
    VX_UP(up_d(A.xb, nullptr, 3, false));
    VX_UP(up_d(A.xw, nullptr, 3, false));
    VX_UP(up_f(A.mass, d->mass, 1, 0));
    VX_UP(up_f(A.young, d->young, 1, 0));
    VX_UP(up_f(A.young_perp, d->young_perp, 1, 0));
    VX_UP(up_f(A.shear, d->shear, 1, 0));
    VX_UP(up_f(A.solid, d->solid, 1, 1));
    VX_UP(up_f(A.fiber, d->fibers, 3, 0));
    VX_UP(up_f(A.aniso, d->anisotropy, 1, 0));
    VX_UP(up_f(A.dtemp, d->delta_temp, 1, 0));
    VX_UP(up_f(A.a_par, d->alpha_par, 1, 0));
    VX_UP(up_f(A.a_perp, d->alpha_perp, 1, 0));
    VX_UP(up_f(A.loads, d->loads, 3, 0));
    VX_UP(up_bond(A.damage, d->damage, 0.0f));
    VX_UP(up_bond(A.kscale, d->kscale, 1.0f));
    VX_UP(up_bond_vec(A.offset, d->formation_offset));
    VX_UP(up_bond_vec(A.plastic, d->plastic_strain));
    VX_UP(up_f(A.strength, d->strength, kStrength, 0.0f));
    for (int a = 0; a < 3; ++a)
        for (int f = 0; f < kHistory; ++f) {
            for (size_t v = 0; v < n; ++v) fbuf[v] = d->bond_history ? d->bond_history[(a * n + v) * kHistory + f] : 0.0f;
            VX_UP(put(A.hist + (a * kHistory + f) * R + voff, fbuf.data(), 4 * n));
        }
    VX_UP(up_d(A.rot_asm, d->rotations, 4, false));
    VX_UP(put(A.kinematic + voff, d->kinematic, n));
    VX_UP(put(A.bonds + voff, bonds.data(), n));
    VX_UP(put(A.tmask + voff, tmask.data(), n));
    VX_UP(put(A.mask + voff, mask.data(), n));
    // Inner vectors start at zero so every plane a first iteration reads is
    // defined.
    std::fill(fbuf.begin(), fbuf.end(), 0.0f);
    std::fill(buf.begin(), buf.end(), 0.0);
    for (float *p : {A.d, A.r, A.z, A.p0, A.p1, A.ap})
        for (int q = 0; q < 3; ++q) VX_UP(put(p + q * R + voff, fbuf.data(), 4 * n));
    for (double *p : {A.dd, A.rd, A.zd, A.p0d, A.p1d, A.apd})
        for (int q = 0; q < 3; ++q) VX_UP(put(p + q * R + voff, buf.data(), 8 * n));
#undef VX_UP
    return VX_OK;
}

SlotParams device_params(double h, const double *g, const vx_solve_params *s, const vx_fracture_params *f) {
    SlotParams sp{};
    if (f && f->enabled) {
        sp.fracture = 1;
        sp.fcfg = FractureCfg{f->damage_rate,       f->energy_release,    f->fatigue_cycles,
                              f->fatigue_exponent, f->fatigue_threshold, f->plastic_heat_fraction};
    }
    sp.h = h;
    sp.gx = g ? g[0] : 0.0;
    sp.gy = g ? g[1] : 0.0;
    sp.gz = g ? g[2] : 0.0;
    if (s) {
        // Plain FP64: the inner solve goes straight to (half) the outer
        // tolerance and one more sweep confirms the true residual.
        const bool f64 = s->mode == VX_SOLVE_F64;
        sp.tol = s->tolerance;
        sp.inner_tol = f64 ? s->tolerance * 0.5 : s->inner_tolerance;
        sp.max_inner = s->max_inner;
        sp.max_outer = f64 ? 2 : s->max_outer;
    }
    return sp;
}

bool solve_ok(const vx_solve_params &s) {
    return s.max_inner > 0 && s.tolerance > 0.0 && s.mode <= VX_SOLVE_F64 && s.preconditioner <= VX_PRECOND_MULTIGRID &&
           !(s.preconditioner == VX_PRECOND_MULTIGRID && s.mode != VX_SOLVE_MIXED) && s.reserved == 0;
}

int32_t admit(vx_batch *b, const vx_scene_desc *d, const vx_step_params *p, uint32_t steps, bool manual, uint32_t *out_slot) {
    std::vector<uint8_t> bonds, tmask, mask;
    uint32_t active = 0;
    if (int32_t c = check_desc(d, bonds, tmask, mask, &active)) return c;
    if (!manual && (p->substeps == 0 || p->corotation_iterations == 0 || !(p->fixed_dt > 0.0) || !solve_ok(p->solve)))
        return fail(VX_ERR_INVALID_ARGUMENT, "invalid step parameters");
    if (!b->single && p && p->solve.preconditioner == VX_PRECOND_MULTIGRID)
        return fail(VX_ERR_INVALID_ARGUMENT, "the multigrid preconditioner needs a one-scene arena (vx_scene_*)");
    uint32_t slot = b->max_slots;
    for (uint32_t i = 0; i < b->max_slots; ++i)
        if (!b->slots[i].used) {
            slot = i;
            break;
        }
    if (slot == b->max_slots) return fail(VX_ERR_TOO_LARGE, "no free scene slot");
    const uint32_t n = d->nx * d->ny * d->nz;
    const uint32_t need = tiles_for(n);
    const uint32_t t0 = find_tiles(b, need);
    if (t0 == b->cap_tiles) return fail(VX_ERR_TOO_LARGE, "no contiguous free range of %u tiles", need);
    // Nothing of the arena may run while its tables change.
    if (int32_t c = drain(b)) return c;
    const SlotGeo g{d->nx, d->ny, d->nz, n, t0, need, t0 * kBlock, slot};
    if (int32_t c = upload_scene(b, d, g, bonds, tmask, mask)) return c;
    if (p && p->fracture.enabled && !d->strength)
        return fail(VX_ERR_INVALID_ARGUMENT, "the fracture pass needs the scene's strength rows");
    const SlotParams sp = p ? device_params(p->fixed_dt / p->substeps, p->gravity, &p->solve, &p->fracture) : SlotParams{};
    SlotState st{};
    st.done = kConv;
    VX_HIP(hipMemcpy(b->d_geo + slot, &g, sizeof g, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(b->d_prm + slot, &sp, sizeof sp, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(b->d_st + slot, &st, sizeof st, hipMemcpyHostToDevice));
    if (int32_t c = set_tiles(b, t0, need, &g)) return c;
    SlotHost &h = b->slots[slot];
    h = SlotHost{};
    h.used = true;
    h.geo = g;
    h.active = active;
    h.manual = manual;
    h.has_strength = d->strength != nullptr;
    if (p) h.params = *p;
    h.steps = steps;
    h.phase = manual ? P_IDLE : (steps ? P_START : P_DONE);
    b->known[slot] = st;
    b->inflight[slot] = 0;
    *out_slot = slot;
    return VX_OK;
}

// Device parameters of one slot; host copies keep the substep length in
// params.fixed_dt and the gravity for later stage calls.
int32_t set_params(vx_batch *b, uint32_t slot, double h, const double *g, const vx_solve_params *s,
                   const vx_fracture_params *f = nullptr) {
    if (int32_t c = drain(b)) return c;
    vx_step_params &hp = b->slots[slot].params;
    const SlotParams sp = device_params(h, g, s, f ? f : &hp.fracture);
    hp.fixed_dt = h;
    if (g)
        for (int i = 0; i < 3; ++i) hp.gravity[i] = g[i];
    if (s) hp.solve = *s;
    if (f) hp.fracture = *f;
    VX_HIP(hipMemcpy(b->d_prm + slot, &sp, sizeof sp, hipMemcpyHostToDevice));
    return VX_OK;
}

int32_t download_planes(const double *dev, size_t R, size_t voff, size_t n, int comps, double *out) {
    std::vector<double> buf(n);
    for (int q = 0; q < comps; ++q) {
        VX_HIP(hipMemcpy(buf.data(), dev + q * R + voff, 8 * n, hipMemcpyDeviceToHost));
        for (size_t v = 0; v < n; ++v) out[comps * v + q] = buf[v];
    }
    return VX_OK;
}

int32_t scene_ok(vx_scene *sc) { return sc && sc->batch ? VX_OK : fail(VX_ERR_INVALID_ARGUMENT, "null scene"); }

}  // namespace
}  // namespace vx

using namespace vx;

// Drop the captured chunk graphs (their shape changed).
static void reset_graphs(vx_batch *b) {
    for (auto &kv : b->g_inner)
        if (kv.second) (void)hipGraphExecDestroy(kv.second);
    b->g_inner.clear();
}

static int32_t set_schedule(vx_batch *b, uint32_t chunk, uint32_t iters) {
    if (iters > 256) return fail(VX_ERR_INVALID_ARGUMENT, "at most 256 iterations per tick");
    // Each chunk graph alternates the p buffers starting from buffer 0, so a
    // chunk must end on the buffer the next one starts from.
    if (iters & 1) return fail(VX_ERR_INVALID_ARGUMENT, "iterations per tick must be even");
    if (int32_t c = drain(b)) return c;
    if (chunk) {
        b->chunk = chunk;
        b->A.chunk = chunk;
    }
    if (iters) b->iters_per_tick = iters;
    reset_graphs(b);
    return VX_OK;
}

// ---- batch API ----------------------------------------------------------------
extern "C" uint64_t vx_batch_bytes_per_row(void) { return uint64_t(kRowBytes); }

extern "C" int32_t vx_batch_create(uint64_t capacity_rows, uint32_t max_scenes, uint32_t xcd_chunk_tiles, vx_batch **out) {
    clear_error();
    return create_batch(capacity_rows, max_scenes, xcd_chunk_tiles, out);
}

extern "C" int32_t vx_batch_destroy(vx_batch *b) {
    clear_error();
    if (!b) return VX_OK;
    Lock lock;
    (void)drain(b);
    (void)hipStreamSynchronize(context().stream);
    reset_graphs(b);
    for (auto &f : b->fl) {
        if (f.bits) (void)hipHostFree(f.bits);
        if (f.list) (void)hipHostFree(f.list);
        if (f.st) (void)hipHostFree(f.st);
        if (f.done) (void)hipEventDestroy(f.done);
        if (f.frac0) (void)hipEventDestroy(f.frac0);
        if (f.frac1) (void)hipEventDestroy(f.frac1);
    }
    if (b->tool) {
        if (b->tool->alloc) (void)hipFree(b->tool->alloc);
        delete b->tool;
        b->tool = nullptr;
    }
    if (b->mg) {
        if (b->mg->alloc) (void)hipFree(b->mg->alloc);
        delete b->mg;
        b->mg = nullptr;
    }
    if (b->alloc) (void)hipFree(b->alloc);
    delete b;
    return VX_OK;
}

extern "C" int32_t vx_batch_set_schedule(vx_batch *b, uint32_t xcd_chunk_tiles, uint32_t iterations_per_tick) {
    clear_error();
    if (!b) return fail(VX_ERR_INVALID_ARGUMENT, "null batch");
    Lock lock;
    return set_schedule(b, xcd_chunk_tiles, iterations_per_tick);
}

extern "C" int32_t vx_batch_free_rows(vx_batch *b, uint64_t *largest, uint64_t *total) {
    clear_error();
    if (!b) return fail(VX_ERR_INVALID_ARGUMENT, "null batch");
    Lock lock;
    uint64_t run = 0, best = 0, tot = 0;
    for (uint32_t t = 0; t < b->cap_tiles; ++t) {
        if (b->tiles[t].slot == kFreeTile) {
            run += 1;
            tot += 1;
            best = std::max(best, run);
        } else {
            run = 0;
        }
    }
    if (largest) *largest = best * kBlock;
    if (total) *total = tot * kBlock;
    return VX_OK;
}

extern "C" int32_t vx_batch_admit(vx_batch *b, const vx_scene_desc *d, const vx_step_params *p, uint32_t steps, uint32_t *out_slot) {
    clear_error();
    if (!b || !p || !out_slot) return fail(VX_ERR_INVALID_ARGUMENT, "null argument");
    Lock lock;
    return admit(b, d, p, steps, false, out_slot);
}

// One scheduler tick with one tick of lookahead: the new tick is decided from
// the states read back after the tick before the one in flight, and enqueued
// before that one is waited for. A scene whose tick in flight is a pure
// inner chunk continues with another chunk (which does nothing if its solve
// has ended in the meantime); a scene whose tick in flight changes its phase
// sits this tick out.
constexpr uint32_t kChunkBits = B_INNER | B_F64 | B_RB | B_MG;  // bits of a pure inner chunk
extern "C" int32_t vx_batch_tick(vx_batch *b, uint32_t *out_running) {
    clear_error();
    if (!b) return fail(VX_ERR_INVALID_ARGUMENT, "null batch");
    Lock lock;
    Flight &f = b->fl[b->cur];
    Flight &prev = b->fl[b->cur ^ 1];
    uint32_t launched = 0;
    for (uint32_t i = 0; i < b->max_slots; ++i) {
        const SlotHost &h = b->slots[i];
        uint32_t bits = 0;
        if (h.used && !h.manual && !h.held && !(b->inflight[i] & ~kChunkBits)) bits = next_bits(h, b->known[i]);
        f.bits[i] = bits;
        launched += bits ? 1 : 0;
    }
    if (launched)
        if (int32_t c = launch_tick(b, f)) return c;
    if (int32_t c = retire(b, prev)) return c;
    for (uint32_t i = 0; i < b->max_slots; ++i) b->inflight[i] = launched ? f.bits[i] : 0u;
    b->cur ^= 1;
    uint32_t still = 0;
    for (uint32_t i = 0; i < b->max_slots; ++i) {
        const SlotHost &h = b->slots[i];
        if (h.used && !h.manual && h.phase != P_DONE && h.phase != P_FAILED) still += 1;
    }
    if (!still)
        if (int32_t c = drain(b)) return c;
    if (out_running) *out_running = still;
    return VX_OK;
}

extern "C" int32_t vx_batch_report(vx_batch *b, uint32_t slot, vx_batch_scene_report *r) {
    clear_error();
    if (!b || !r || slot >= b->max_slots) return fail(VX_ERR_INVALID_ARGUMENT, "invalid slot");
    Lock lock;
    const SlotHost &h = b->slots[slot];
    std::memset(r, 0, sizeof *r);
    if (!h.used) return VX_OK;
    r->state = h.phase == P_DONE ? VX_SCENE_DONE : (h.phase == P_FAILED ? VX_SCENE_FAILED : VX_SCENE_RUNNING);
    r->status = h.status;
    r->steps_done = h.steps_done;
    r->solves = h.solves;
    r->inner_iterations = h.inner;
    r->outer_sweeps = h.sweeps;
    r->rows = h.geo.n;
    r->active_rows = h.active;
    r->last_relative_residual = h.last_rel;
    r->last_angle = h.last_angle;
    r->fracture = h.frac;
    return VX_OK;
}

extern "C" int32_t vx_batch_states(vx_batch *b, uint32_t *states) {
    clear_error();
    if (!b || !states) return fail(VX_ERR_INVALID_ARGUMENT, "null argument");
    Lock lock;
    for (uint32_t i = 0; i < b->max_slots; ++i) {
        const SlotHost &h = b->slots[i];
        states[i] = !h.used ? VX_SCENE_FREE
                            : (h.phase == P_DONE ? VX_SCENE_DONE : (h.phase == P_FAILED ? VX_SCENE_FAILED : VX_SCENE_RUNNING));
    }
    return VX_OK;
}

extern "C" int32_t vx_batch_download(vx_batch *b, uint32_t slot, double *positions, double *rotations, double *velocities) {
    clear_error();
    if (!b || slot >= b->max_slots || !b->slots[slot].used) return fail(VX_ERR_INVALID_ARGUMENT, "invalid slot");
    Lock lock;
    if (int32_t c = drain(b)) return c;
    const SlotGeo &g = b->slots[slot].geo;
    const size_t R = b->A.R;
    if (positions)
        if (int32_t c = download_planes(b->A.pos, R, g.voff, g.n, 3, positions)) return c;
    if (rotations)
        if (int32_t c = download_planes(b->A.rot, R, g.voff, g.n, 4, rotations)) return c;
    if (velocities)
        if (int32_t c = download_planes(b->A.vel, R, g.voff, g.n, 3, velocities)) return c;
    return VX_OK;
}

// Bond state of a slot: neighbour from the bond bits, damage, plastic strain
// and history planes, re-laid out per bond as [axis][voxel][...].
static int32_t download_bonds(vx_batch *b, uint32_t slot, uint32_t *neighbor, float *damage, float *plastic, float *history) {
    if (int32_t c = drain(b)) return c;
    const SlotGeo &g = b->slots[slot].geo;
    const DevArena &A = b->A;
    const size_t R = A.R, n = g.n, voff = g.voff;
    std::vector<float> buf(n);
    auto get = [&](const float *plane) -> int32_t {
        VX_HIP(hipMemcpy(buf.data(), plane + voff, 4 * n, hipMemcpyDeviceToHost));
        return VX_OK;
    };
    if (neighbor) {
        std::vector<uint8_t> bits(n);
        VX_HIP(hipMemcpy(bits.data(), A.bonds + voff, n, hipMemcpyDeviceToHost));
        for (int a = 0; a < 3; ++a) {
            const uint32_t stride = a == 0 ? 1u : (a == 1 ? g.nx : g.nx * g.ny);
            for (size_t v = 0; v < n; ++v) neighbor[a * n + v] = (bits[v] >> a & 1) ? uint32_t(v) + stride : VX_EMPTY;
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (damage) {
            if (int32_t c = get(A.damage + a * R)) return c;
            std::memcpy(damage + a * n, buf.data(), 4 * n);
        }
        for (int q = 0; plastic && q < 3; ++q) {
            if (int32_t c = get(A.plastic + (a * 3 + q) * R)) return c;
            for (size_t v = 0; v < n; ++v) plastic[(a * n + v) * 3 + q] = buf[v];
        }
        for (int f = 0; history && f < kHistory; ++f) {
            if (int32_t c = get(A.hist + (a * kHistory + f) * R)) return c;
            for (size_t v = 0; v < n; ++v) history[(a * n + v) * kHistory + f] = buf[v];
        }
    }
    return VX_OK;
}

extern "C" int32_t vx_batch_download_bonds(vx_batch *b, uint32_t slot, uint32_t *neighbor, float *damage, float *plastic_strain,
                                           float *history) {
    clear_error();
    if (!b || slot >= b->max_slots || !b->slots[slot].used) return fail(VX_ERR_INVALID_ARGUMENT, "invalid slot");
    Lock lock;
    return download_bonds(b, slot, neighbor, damage, plastic_strain, history);
}

// Each bond's last force and multiplier of the slot's last substep, (3, n, 3)
// [axis][voxel][component], f64.
static int32_t download_bond_forces(vx_batch *b, uint32_t slot, double *last_force, double *lambda) {
    if (int32_t c = drain(b)) return c;
    const SlotGeo &g = b->slots[slot].geo;
    const DevArena &A = b->A;
    const size_t R = A.R, n = g.n;
    for (int a = 0; a < 3; ++a) {
        if (last_force)
            if (int32_t c = download_planes(A.force + size_t(a) * 3 * R, R, g.voff, n, 3, last_force + size_t(a) * n * 3)) return c;
        if (lambda)
            if (int32_t c = download_planes(A.lambda + size_t(a) * 3 * R, R, g.voff, n, 3, lambda + size_t(a) * n * 3)) return c;
    }
    return VX_OK;
}

extern "C" int32_t vx_batch_download_bond_forces(vx_batch *b, uint32_t slot, double *last_force, double *lambda) {
    clear_error();
    if (!b || slot >= b->max_slots || !b->slots[slot].used) return fail(VX_ERR_INVALID_ARGUMENT, "invalid slot");
    Lock lock;
    return download_bond_forces(b, slot, last_force, lambda);
}

extern "C" int32_t vx_batch_hold(vx_batch *b, uint32_t slot, uint32_t hold) {
    clear_error();
    if (!b || slot >= b->max_slots) return fail(VX_ERR_INVALID_ARGUMENT, "invalid slot");
    Lock lock;
    SlotHost &h = b->slots[slot];
    if (!h.used) return fail(VX_ERR_INVALID_ARGUMENT, "slot %u holds no scene", slot);
    h.held = hold != 0;
    return VX_OK;
}

extern "C" int32_t vx_batch_release(vx_batch *b, uint32_t slot) {
    clear_error();
    if (!b || slot >= b->max_slots) return fail(VX_ERR_INVALID_ARGUMENT, "invalid slot");
    Lock lock;
    SlotHost &h = b->slots[slot];
    if (!h.used) return VX_OK;
    if (int32_t c = drain(b)) return c;
    if (int32_t c = set_tiles(b, h.geo.tile0, h.geo.ntiles, nullptr)) return c;
    h = SlotHost{};
    const SlotGeo g{};
    VX_HIP(hipMemcpy(b->d_geo + slot, &g, sizeof g, hipMemcpyHostToDevice));
    return VX_OK;
}

// ---- single-scene API (a one-scene arena driven stage by stage) ------------------
extern "C" int32_t vx_scene_create(const vx_scene_desc *desc, vx_scene **out) {
    clear_error();
    if (!desc || !out) return fail(VX_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint64_t n = uint64_t(desc->nx) * desc->ny * desc->nz;
    if (n == 0) return fail(VX_ERR_INVALID_ARGUMENT, "empty scene");
    if (n >= (1ull << 31)) return fail(VX_ERR_TOO_LARGE, "scene of %llu REVs exceeds the index range", (unsigned long long)n);
    Lock lock;
    vx_batch *b = nullptr;
    if (int32_t c = create_batch(n, 1, 0, &b)) return c;
    uint32_t slot = 0;
    if (int32_t c = admit(b, desc, nullptr, 0, true, &slot)) {
        vx_batch_destroy(b);
        return c;
    }
    b->single = true;
    vx_scene *sc = new vx_scene();
    sc->batch = b;
    sc->slot = slot;
    sc->active = b->slots[slot].active;
    *out = sc;
    return VX_OK;
}

extern "C" int32_t vx_scene_set_schedule(vx_scene *sc, uint32_t iterations_per_tick) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    return set_schedule(sc->batch, 0, iterations_per_tick);
}

extern "C" int32_t vx_scene_set_shape(vx_scene *sc, uint32_t zcols, uint32_t deal, uint32_t update_threads) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (zcols > 64 || (zcols & (zcols - 1)))
        return fail(VX_ERR_INVALID_ARGUMENT, "z-column group %u is not 0 or a power of two up to 64", zcols);
    if (deal > VX_DEAL_ROUND_ROBIN) return fail(VX_ERR_INVALID_ARGUMENT, "unknown XCD deal %u", deal);
    if (update_threads != kBlock && update_threads != kHalf)
        return fail(VX_ERR_INVALID_ARGUMENT, "update threads per tile %u is not %u or %u", update_threads, kBlock, kHalf);
    Lock lock;
    vx_batch *b = sc->batch;
    if (int32_t c = drain(b)) return c;
    b->zcols = zcols;
    b->deal = deal;
    b->update_threads = update_threads;
    reset_graphs(b);
    return VX_OK;
}

extern "C" int32_t vx_scene_destroy(vx_scene *sc) {
    clear_error();
    if (!sc) return VX_OK;
    vx_batch_destroy(sc->batch);
    delete sc;
    return VX_OK;
}

// ---- rigid tool ------------------------------------------------------------------
static bool finite3(const double *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

extern "C" int32_t vx_scene_set_tool(vx_scene *sc, const vx_tool_desc *d) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    vx_batch *b = sc->batch;
    if (int32_t c = drain(b)) return c;
    if (b->tool) {
        VX_HIP(hipStreamSynchronize(context().stream));
        if (b->tool->alloc) VX_HIP(hipFree(b->tool->alloc));
        delete b->tool;
        b->tool = nullptr;
    }
    b->slots[sc->slot].has_tool = false;
    if (!d) return VX_OK;
    if (d->n_cells == 0 || !d->cells || !(d->mass > 0.0) || !std::isfinite(d->mass) || !finite3(d->inverse_inertia) ||
        !finite3(d->position) || !finite3(d->velocity) || !finite3(d->angular_velocity) || !finite3(d->external_force) ||
        !finite3(d->external_torque) || !std::isfinite(d->friction) || !(d->friction >= 0.0))
        return fail(VX_ERR_INVALID_ARGUMENT, "invalid tool description");
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(d->rotation[k])) return fail(VX_ERR_INVALID_ARGUMENT, "invalid tool rotation");
    if (d->has_mount) {
        // the guide direction is a unit vector or zero (a base fixed in place), as mount_base makes it
        const double dd = (d->mount_direction[0] * d->mount_direction[0] + d->mount_direction[1] * d->mount_direction[1]) +
                          d->mount_direction[2] * d->mount_direction[2];
        bool ok = finite3(d->mount_local_base) && finite3(d->mount_world_base) && finite3(d->mount_direction) &&
                  (dd == 0.0 || std::fabs(dd - 1.0) <= 1e-6);
        for (int k = 0; k < 4; ++k) ok = ok && std::isfinite(d->mount_rotation[k]);
        if (!ok) return fail(VX_ERR_INVALID_ARGUMENT, "invalid tool mount");
    }
    const uint32_t nc = d->n_cells, ntiles = b->slots[sc->slot].geo.ntiles;
    std::vector<double> cells(3 * size_t(nc));
    double radius = 0.0;
    ToolDev T{};
    for (int k = 0; k < 3; ++k) {
        T.cell_lo[k] = INFINITY;
        T.cell_hi[k] = -INFINITY;
    }
    for (uint32_t c = 0; c < nc; ++c) {
        for (int k = 0; k < 3; ++k) {
            cells[3 * c + k] = double(d->cells[3 * c + k]);
            if (!std::isfinite(cells[3 * c + k])) return fail(VX_ERR_INVALID_ARGUMENT, "non-finite tool cell");
            T.cell_lo[k] = std::min(T.cell_lo[k], cells[3 * c + k]);
            T.cell_hi[k] = std::max(T.cell_hi[k], cells[3 * c + k]);
        }
        radius = std::max(radius, std::sqrt((cells[3 * c] * cells[3 * c] + cells[3 * c + 1] * cells[3 * c + 1]) +
                                            cells[3 * c + 2] * cells[3 * c + 2]));
    }
    T.mass = d->mass;
    T.friction = d->friction;
    T.radius = radius;
    for (int k = 0; k < 3; ++k) {
        T.inv_inertia[k] = d->inverse_inertia[k];
        T.pos[k] = d->position[k];
        T.vel[k] = d->velocity[k];
        T.ang[k] = d->angular_velocity[k];
        T.force[k] = d->external_force[k];
        T.torque[k] = d->external_torque[k];
        T.m_local[k] = d->mount_local_base[k];
        T.m_world[k] = d->mount_world_base[k];
        T.m_dir[k] = d->mount_direction[k];
    }
    for (int k = 0; k < 4; ++k) {
        T.rot[k] = d->rotation[k];
        T.m_rot[k] = d->mount_rotation[k];
    }
    T.n_cells = nc;
    T.has_mount = d->has_mount ? 1u : 0u;
    // One allocation, carved like the arena.
    ToolHost *th = new ToolHost();
    ToolBufs &tb = th->bufs;
    struct Item {
        void **ptr;
        size_t bytes;
    } items[] = {
        {(void **)&tb.dev, sizeof(ToolDev)}, {(void **)&tb.cells, 24 * size_t(nc)},
        {(void **)&tb.tile_count, 4 * size_t(ntiles)}, {(void **)&tb.tile_off, 4 * size_t(ntiles)},
        {(void **)&tb.key, 8 * size_t(kToolMaxCand)}, {(void **)&tb.key2, 8 * size_t(kToolMaxCand)},
        {(void **)&tb.idx, 4 * size_t(kToolMaxCand)}, {(void **)&tb.idx2, 4 * size_t(kToolMaxCand)},
        {(void **)&tb.cv, 24 * size_t(kToolMaxCand)}, {(void **)&tb.con, sizeof(ToolContact) * kToolCellSlots * size_t(nc)},
        {(void **)&tb.ncon, 4 * size_t(nc)}, {(void **)&tb.vsave, 24 * size_t(kToolMaxCand)},
    };
    size_t total = 0;
    for (auto &it : items) total += (it.bytes + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&th->alloc, total);
    if (e != hipSuccess) {
        delete th;
        return fail(VX_ERR_TOO_LARGE, "tool buffers of %zu bytes: %s", total, hipGetErrorString(e));
    }
    size_t off = 0;
    for (auto &it : items) {
        *it.ptr = static_cast<char *>(th->alloc) + off;
        off += (it.bytes + 255) & ~size_t(255);
    }
    th->n_cells = nc;
    b->tool = th;
    VX_HIP(hipMemcpy(tb.dev, &T, sizeof T, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(const_cast<double *>(tb.cells), cells.data(), 24 * size_t(nc), hipMemcpyHostToDevice));
    b->slots[sc->slot].has_tool = true;
    return VX_OK;
}

// The tool's device state after the last advance.
static int32_t read_tool(vx_batch *b, ToolDev &T) {
    if (!b->tool) return fail(VX_ERR_INVALID_ARGUMENT, "the scene has no tool");
    if (int32_t c = drain(b)) return c;
    VX_HIP(hipMemcpy(&T, b->tool->bufs.dev, sizeof T, hipMemcpyDeviceToHost));
    return VX_OK;
}

static int32_t tool_status(const ToolDev &T) {
    if (T.status & TS_CAND)
        return fail(VX_ERR_TOO_LARGE, "the tool's neighbourhood holds %u REVs, more than %u", T.candidates, kToolMaxCand);
    if (T.status & TS_SLOTS) return fail(VX_ERR_TOO_LARGE, "a tool cell touched more than %u REVs", kToolCellSlots);
    if (T.status & TS_KEY) return fail(VX_ERR_INVALID_ARGUMENT, "a REV near the tool lies outside the bin range");
    if (T.status & TS_MARGIN) return fail(VX_ERR_DEVICE, "the tool moved beyond its candidate margin in one substep");
    if (T.status & TS_STEPS)
        return fail(VX_ERR_INVALID_ARGUMENT, "the tool would need more than %u microsteps in one substep", kToolMaxSteps);
    return VX_OK;
}

extern "C" int32_t vx_scene_tool_state(vx_scene *sc, vx_tool_state *out) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (!out) return fail(VX_ERR_INVALID_ARGUMENT, "null state");
    Lock lock;
    ToolDev T{};
    if (int32_t c = read_tool(sc->batch, T)) return c;
    for (int k = 0; k < 3; ++k) {
        out->position[k] = T.pos[k];
        out->velocity[k] = T.vel[k];
        out->angular_velocity[k] = T.ang[k];
    }
    for (int k = 0; k < 4; ++k) out->rotation[k] = T.rot[k];
    out->friction_heat = T.friction_heat;
    out->microsteps = T.steps;
    out->contacts = T.contacts;
    out->candidates = T.candidates;
    out->status = T.status;
    return VX_OK;
}

extern "C" int32_t vx_scene_predict(vx_scene *sc, double h, const double *g) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (!(h > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "substep length must be positive");
    Lock lock;
    vx_batch *b = sc->batch;
    if (int32_t c = set_params(b, sc->slot, h, g, &b->slots[sc->slot].params.solve)) return c;
    return manual_tick(b, sc->slot, B_PREDICT);
}

extern "C" int32_t vx_scene_fit_rotations(vx_scene *sc, double *out_max_angle) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    if (int32_t c = manual_tick(sc->batch, sc->slot, B_FITB)) return c;
    if (out_max_angle) *out_max_angle = sc->batch->known[sc->slot].angle;
    return VX_OK;
}

extern "C" int32_t vx_scene_assemble(vx_scene *sc, double h) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (!(h > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "substep length must be positive");
    Lock lock;
    vx_batch *b = sc->batch;
    const vx_step_params p = b->slots[sc->slot].params;
    if (int32_t c = set_params(b, sc->slot, h, p.gravity, &p.solve)) return c;
    return manual_tick(b, sc->slot, B_ASSEMBLE);
}

// One solve of a vx_scene (the caller holds the context lock); with a timer,
// every tick's timed dispatches are harvested once the tick has retired.
static int32_t scene_solve(vx_scene *sc, const vx_solve_params *prm, vx_solve_report *rep, DispatchTimer *timer) {
    vx_batch *b = sc->batch;
    SlotHost &h = b->slots[sc->slot];
    const vx_step_params p = h.params;
    if (int32_t c = set_params(b, sc->slot, p.fixed_dt, p.gravity, prm)) return c;
    const uint32_t prec = solve_bits(*prm);
    uint32_t bits = B_RESID0 | B_INNER | prec | ((prm->warm_start && h.warm_ok) ? uint32_t(B_WARM_USE) : 0u);
    if (rep) std::memset(rep, 0, sizeof *rep);
    const SlotState *st = &b->known[sc->slot];
    for (;;) {
        if (int32_t c = manual_tick(b, sc->slot, bits)) return c;
        if (timer)
            if (int32_t c = timer->harvest()) return c;
        if (st->nonfinite || st->done == kNonFin) {
            if (rep) rep->status = VX_ERR_NONFINITE_RESIDUAL;
            return fail(VX_ERR_NONFINITE_RESIDUAL, "HIP PCG produced a non-finite residual norm");
        }
        if (st->outer_done) break;
        bits = B_INNER | prec;
    }
    // The solution becomes the warm start of a later solve of this scene.
    if (int32_t c = manual_tick(b, sc->slot, B_WARM_SAVE)) return c;
    h.warm_ok = true;
    const double rel = st->rel;
    if (rep) {
        rep->inner_iterations = st->inner_total;
        rep->outer_sweeps = st->sweeps;
        rep->active_rows = sc->active;
        rep->relative_residual = rel;
        rep->rhs_norm = std::sqrt(st->bb);
    }
    int32_t status = VX_OK;
    if (!(rel <= prm->tolerance))
        status = fail(VX_ERR_NOT_CONVERGED, "HIP PCG did not converge: inner iterations=%u, sweeps=%u, relative residual=%.3e, tolerance=%.3e",
                      st->inner_total, st->sweeps, rel, prm->tolerance);
    if (rep) rep->status = uint32_t(status);
    return status;
}

extern "C" int32_t vx_scene_solve(vx_scene *sc, const vx_solve_params *prm, vx_solve_report *rep) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (!prm || !solve_ok(*prm)) return fail(VX_ERR_INVALID_ARGUMENT, "invalid solve parameters");
    Lock lock;
    return scene_solve(sc, prm, rep, nullptr);
}

extern "C" int32_t vx_scene_solve_timed(vx_scene *sc, const vx_solve_params *prm, vx_solve_report *rep,
                                        vx_kernel_times *times) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (!prm || !solve_ok(*prm) || !times) return fail(VX_ERR_INVALID_ARGUMENT, "invalid solve parameters");
    Lock lock;
    if (int32_t c = drain(sc->batch)) return c;
    DispatchTimer timer;
    g_timer = &timer;
    const int32_t status = scene_solve(sc, prm, rep, &timer);
    g_timer = nullptr;
    // A failed tick may leave timed dispatches unharvested: wait for them.
    (void)hipStreamSynchronize(context().stream);
    if (timer.used) (void)timer.harvest();
    std::memset(times, 0, sizeof *times);
    for (int k = 0; k < DispatchTimer::kKinds; ++k) {
        std::vector<float> &u = timer.us[k];
        vx_kernel_stats &o = k == 0 ? times->spmv : times->update;
        o.dispatches = uint32_t(u.size());
        if (u.empty()) continue;
        double sum = 0.0;
        for (float x : u) sum += x;
        std::sort(u.begin(), u.end());
        o.mean_us = sum / double(u.size());
        o.median_us = u[u.size() / 2];
        o.min_us = u.front();
        o.max_us = u.back();
        o.total_us = sum;
    }
    return status;
}

extern "C" int32_t vx_scene_apply(vx_scene *sc) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    return manual_tick(sc->batch, sc->slot, B_APPLY);
}

extern "C" int32_t vx_scene_finish_substep(vx_scene *sc, double h) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (!(h > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "substep length must be positive");
    Lock lock;
    vx_batch *b = sc->batch;
    const vx_step_params p = b->slots[sc->slot].params;
    if (int32_t c = set_params(b, sc->slot, h, p.gravity, &p.solve)) return c;
    return manual_tick(b, sc->slot, B_FINISH);
}

extern "C" int32_t vx_scene_download(vx_scene *sc, double *positions, double *rotations, double *velocities, double *correction) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    vx_batch *b = sc->batch;
    if (int32_t c = vx_batch_download(b, sc->slot, positions, rotations, velocities)) return c;
    if (correction) {
        const SlotGeo &g = b->slots[sc->slot].geo;
        const double *x = b->known[sc->slot].xpar ? b->A.xb : b->A.xa;
        if (int32_t c = download_planes(x, b->A.R, g.voff, g.n, 3, correction)) return c;
    }
    return VX_OK;
}

extern "C" int32_t vx_scene_download_bond_forces(vx_scene *sc, double *last_force, double *lambda) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    return download_bond_forces(sc->batch, sc->slot, last_force, lambda);
}

extern "C" int32_t vx_scene_download_bonds(vx_scene *sc, uint32_t *neighbor, float *damage, float *plastic_strain, float *history) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    return download_bonds(sc->batch, sc->slot, neighbor, damage, plastic_strain, history);
}

// Replace the scene's axis targets and applied loads for the following steps
// (the per-step inputs of step_with_axis_targets_and_loads,
// src/solver.rs:440-482). The warm start is dropped: the solved axes change.
extern "C" int32_t vx_scene_set_targets_and_loads(vx_scene *sc, const float *targets, const float *loads) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    vx_batch *b = sc->batch;
    if (int32_t c = drain(b)) return c;
    SlotHost &h = b->slots[sc->slot];
    const SlotGeo &g = h.geo;
    const DevArena &A = b->A;
    const size_t n = g.n, voff = g.voff, R = A.R;
    std::vector<uint8_t> kin(n), tmask, mask;
    VX_HIP(hipMemcpy(kin.data(), A.kinematic + voff, n, hipMemcpyDeviceToHost));
    uint32_t active = 0;
    if (int32_t c = free_masks(kin.data(), targets, uint32_t(n), tmask, mask, &active)) return c;
    std::vector<double> buf(n);
    std::vector<float> fbuf(n);
    for (int q = 0; q < 3; ++q) {
        for (size_t v = 0; v < n; ++v) {
            const float t = targets ? targets[3 * v + q] : NAN;
            buf[v] = std::isnan(t) ? 0.0 : double(t);
            fbuf[v] = loads ? loads[3 * v + q] : 0.0f;
        }
        if (int32_t c = put(A.targets + q * R + voff, buf.data(), 8 * n)) return c;
        if (int32_t c = put(A.loads + q * R + voff, fbuf.data(), 4 * n)) return c;
    }
    if (int32_t c = put(A.tmask + voff, tmask.data(), n)) return c;
    if (int32_t c = put(A.mask + voff, mask.data(), n)) return c;
    h.active = active;
    sc->active = active;
    h.warm_ok = false;
    return VX_OK;
}

// Reference row packing (src/solver.rs:1652-1672, 1731-1777) of the assembled
// structured operator: slots in incidence order (-x, +x, -y, +y, -z, +z with
// absent bonds compacted out), off-diagonal blocks -P_i k P_j, zero blocks and
// EMPTY for inactive neighbours.
extern "C" int32_t vx_scene_multigrid_level(vx_scene *sc, uint32_t level, uint32_t *dims, float *diagonal,
                                            float *couplings, uint8_t *masks) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    vx_batch *b = sc->batch;
    if (!b->mg || b->mg->built == ~0ull || level == 0)
        return fail(VX_ERR_INVALID_ARGUMENT, "no multigrid hierarchy (solve with VX_PRECOND_MULTIGRID first) or level 0");
    if (int32_t c = drain(b)) return c;
    VX_HIP(hipStreamSynchronize(context().stream));
    if (level > b->mg->lv.size()) {
        if (dims) dims[0] = dims[1] = dims[2] = 0;
        return VX_OK;
    }
    const MgLevel &L = b->mg->lv[level - 1];
    if (dims) {
        dims[0] = L.nx;
        dims[1] = L.ny;
        dims[2] = L.nz;
    }
    const size_t n = L.n;
    auto node_major = [&](const float *dev, int planes, float *out) -> int32_t {
        std::vector<float> h(size_t(planes) * n);
        VX_HIP(hipMemcpy(h.data(), dev, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i)
            for (int q = 0; q < planes; ++q) out[i * planes + q] = h[size_t(q) * n + i];
        return VX_OK;
    };
    if (diagonal)
        if (int32_t c = node_major(L.D, 6, diagonal)) return c;
    if (couplings)
        if (int32_t c = node_major(L.C, 27, couplings)) return c;
    if (masks) VX_HIP(hipMemcpy(masks, L.mask, n, hipMemcpyDeviceToHost));
    return VX_OK;
}

extern "C" int32_t vx_scene_export_rows(vx_scene *sc, uint32_t *active_out, uint32_t *neighbors, double *blocks, double *diagonal,
                                        double *rhs_out, uint32_t *out_active_count) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    Lock lock;
    vx_batch *b = sc->batch;
    if (int32_t c = drain(b)) return c;
    const SlotGeo &G = b->slots[sc->slot].geo;
    const size_t n = G.n, R = b->A.R, voff = G.voff;
    std::vector<double> k(18 * n), d(6 * n), rhs(3 * n);
    std::vector<uint8_t> mask(n), bonds(n);
    for (int q = 0; q < 18; ++q) VX_HIP(hipMemcpy(k.data() + q * n, b->A.k64 + q * R + voff, 8 * n, hipMemcpyDeviceToHost));
    for (int q = 0; q < 6; ++q) VX_HIP(hipMemcpy(d.data() + q * n, b->A.d64 + q * R + voff, 8 * n, hipMemcpyDeviceToHost));
    for (int q = 0; q < 3; ++q) VX_HIP(hipMemcpy(rhs.data() + q * n, b->A.rhs + q * R + voff, 8 * n, hipMemcpyDeviceToHost));
    VX_HIP(hipMemcpy(mask.data(), b->A.mask + voff, n, hipMemcpyDeviceToHost));
    VX_HIP(hipMemcpy(bonds.data(), b->A.bonds + voff, n, hipMemcpyDeviceToHost));
    std::vector<uint32_t> aidx(n, VX_EMPTY);
    uint32_t na = 0;
    for (size_t v = 0; v < n; ++v)
        if (mask[v]) aidx[v] = na++;
    const uint32_t nx = G.nx, ny = G.ny;
    const int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (size_t v = 0; v < n; ++v) {
        if (!mask[v]) continue;
        const uint32_t slot = aidx[v];
        if (active_out) active_out[slot] = uint32_t(v);
        if (diagonal)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) diagonal[9 * size_t(slot) + 3 * r + c] = d[sym[r][c] * n + v];
        if (rhs_out)
            for (int q = 0; q < 3; ++q) rhs_out[3 * size_t(slot) + q] = rhs[q * n + v];
        if (neighbors)
            for (int s = 0; s < 6; ++s) neighbors[6 * size_t(slot) + s] = VX_EMPTY;
        if (blocks)
            for (int s = 0; s < 54; ++s) blocks[54 * size_t(slot) + s] = 0.0;
        uint32_t e = 0;
        const uint32_t c3[3] = {uint32_t(v % nx), uint32_t((v / nx) % ny), uint32_t(v / (size_t(nx) * ny))};
        for (int a = 0; a < 3; ++a) {
            const size_t s = a == 0 ? 1 : (a == 1 ? nx : size_t(nx) * ny);
            for (int side = 0; side < 2; ++side) {
                size_t owner, other;
                if (side == 0) {
                    if (c3[a] == 0) continue;
                    owner = v - s;
                    other = v - s;
                } else {
                    owner = v;
                    other = v + s;
                }
                if (!(bonds[owner] >> a & 1)) continue;
                if (!(mask[v] | mask[other])) continue;
                if (e >= 6) continue;
                const uint32_t nb = aidx[other];
                if (neighbors) neighbors[6 * size_t(slot) + e] = nb;
                if (blocks && nb != VX_EMPTY)
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            const double kv = k[(size_t(a) * 6 + sym[r][c]) * n + owner];
                            const double pr = (mask[v] >> r) & 1, pc = (mask[other] >> c) & 1;
                            blocks[54 * size_t(slot) + 9 * e + 3 * r + c] = -pr * kv * pc;
                        }
                e += 1;
            }
        }
    }
    if (out_active_count) *out_active_count = na;
    return VX_OK;
}

extern "C" int32_t vx_scene_time_spmv(vx_scene *sc, uint32_t reps, double *out_us) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (reps == 0 || !out_us) return fail(VX_ERR_INVALID_ARGUMENT, "reps must be positive");
    Lock lock;
    vx_batch *b = sc->batch;
    if (int32_t c = drain(b)) return c;
    Context &ctx = context();
    // The timed launches run as first FP32 inner iterations of the scene (k = 0,
    // unit residual partials, never done); they keep the solution buffers and
    // leave the inner vectors' contents unspecified.
    SlotState st = b->known[sc->slot];
    st.k_spmv = 0;
    st.k_update = 0;
    st.seg = 0;
    st.iters = 0;
    st.done = kRun;
    st.red[kRz] = st.red[kRr] = double(b->slots[sc->slot].geo.ntiles);
    st.rho[0] = st.rho[1] = 0.0;
    const vx_step_params p = b->slots[sc->slot].params;
    SlotParams sp = device_params(p.fixed_dt, p.gravity, &p.solve, &p.fracture);
    sp.inner_tol = 0.0;
    sp.max_inner = 1u << 30;
    const SlotGeo &g = b->slots[sc->slot].geo;
    VX_HIP(hipMemcpy(b->d_st + sc->slot, &st, sizeof st, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(b->d_prm + sc->slot, &sp, sizeof sp, hipMemcpyHostToDevice));
    const std::vector<double> ones(g.ntiles, 1.0);
    VX_HIP(hipMemcpy(b->A.partials + size_t(kRz) * b->cap_tiles + g.tile0, ones.data(), 8 * g.ntiles, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(b->A.partials + size_t(kRr) * b->cap_tiles + g.tile0, ones.data(), 8 * g.ntiles, hipMemcpyHostToDevice));
    std::vector<uint32_t> bits(b->max_slots, 0u);
    bits[sc->slot] = B_INNER;
    VX_HIP(hipMemcpy(b->d_bits, bits.data(), sizeof(uint32_t) * b->max_slots, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(spread_bits, dim3((b->cap_tiles + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx.stream, b->A, BitsArg{});
    b->spread_dirty = true;
    VX_HIP(hipGetLastError());
    const OneScene one = make_one(g, deal_spmv_rr(b), b->zcols);
    const bool early = g.ntiles > kBigTiles;  // as the solve's chunks launch it
    // The launches run back to back inside one graph, as the solve runs them,
    // between two events on the library stream.
    auto enqueue = [&]() {
        for (uint32_t k = 0; k < reps; ++k)
            if (early)
                hipLaunchKernelGGL((pcg_bond_spmv<float, true, false, true>), dim3(one.grid), dim3(kBlock), 0, ctx.stream, b->A,
                                   one, int(k & 1));
            else
                hipLaunchKernelGGL((pcg_bond_spmv<float, true, false, false>), dim3(one.grid), dim3(kBlock), 0, ctx.stream, b->A,
                                   one, int(k & 1));
    };
    hipGraphExec_t exec = nullptr;
    if (!ctx.direct_launch) {
        hipGraph_t graph;
        VX_HIP(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
        enqueue();
        VX_HIP(hipStreamEndCapture(ctx.stream, &graph));
        hipError_t ge = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ge != hipSuccess) return fail(VX_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(ge));
    }
    auto replay = [&]() -> hipError_t {
        if (exec) return hipGraphLaunch(exec, ctx.stream);
        enqueue();
        return hipGetLastError();
    };
    // Five timed replays after a warm-up one; the median replay is reported.
    constexpr int kReplays = 5;
    hipEvent_t ev[kReplays + 1];
    for (auto &e : ev) VX_HIP(hipEventCreate(&e));
    VX_HIP(replay());  // warm-up
    VX_HIP(hipEventRecord(ev[0], ctx.stream));
    for (int i = 0; i < kReplays; ++i) {
        VX_HIP(replay());
        VX_HIP(hipEventRecord(ev[i + 1], ctx.stream));
    }
    VX_HIP(hipStreamSynchronize(ctx.stream));
    float ms[kReplays];
    for (int i = 0; i < kReplays; ++i) VX_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    std::sort(ms, ms + kReplays);
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (exec) (void)hipGraphExecDestroy(exec);
    *out_us = 1000.0 * double(ms[kReplays / 2]) / reps;
    // Restore the scene's state word (solution buffer parity) and parameters.
    st = b->known[sc->slot];
    VX_HIP(hipMemcpy(b->d_st + sc->slot, &st, sizeof st, hipMemcpyHostToDevice));
    return set_params(b, sc->slot, p.fixed_dt, p.gravity, &p.solve);
}

extern "C" int32_t vx_scene_step(vx_scene *sc, const vx_step_params *prm, vx_step_report *rep) {
    clear_error();
    if (int32_t c = scene_ok(sc)) return c;
    if (!prm || prm->substeps == 0 || prm->corotation_iterations == 0 || !(prm->fixed_dt > 0.0) || !solve_ok(prm->solve))
        return fail(VX_ERR_INVALID_ARGUMENT, "invalid step parameters");
    Lock lock;
    vx_batch *b = sc->batch;
    SlotHost &h = b->slots[sc->slot];
    if (prm->fracture.enabled && !h.has_strength)
        return fail(VX_ERR_INVALID_ARGUMENT, "the fracture pass needs the scene's strength rows");
    const auto t0 = std::chrono::steady_clock::now();
    // Drive the slot through one step with the per-scene state machine.
    if (int32_t c = set_params(b, sc->slot, prm->fixed_dt / prm->substeps, prm->gravity, &prm->solve, &prm->fracture)) return c;
    const bool warm_ok = h.warm_ok;
    h.params = *prm;
    h.manual = false;
    h.steps = 1;
    h.steps_done = 0;
    h.substep = 0;
    h.pass = 0;
    h.finish_next = false;
    h.solves = 0;
    h.inner = 0;
    h.sweeps = 0;
    h.frac = vx_fracture_report{};
    h.status = VX_OK;
    h.warm_ok = warm_ok;
    h.phase = P_START;
    if (b->tool) {
        hipLaunchKernelGGL(tool_reset, dim3(1), dim3(1), 0, context().stream, b->tool->bufs);
        VX_HIP(hipGetLastError());
    }
    uint32_t running = 1;
    int32_t status = VX_OK;
    while (running) {
        if ((status = vx_batch_tick(b, &running))) break;
    }
    (void)drain(b);
    h.manual = true;
    if (status == VX_OK && b->tool) {
        ToolDev T{};
        if ((status = read_tool(b, T)) == VX_OK) status = tool_status(T);
    }
    if (status == VX_OK && h.phase == P_FAILED)
        status = fail(int32_t(h.status), h.status == VX_ERR_NONFINITE_RESIDUAL ? "HIP PCG produced a non-finite residual norm"
                                                                                 : "HIP PCG did not converge within the refinement sweeps");
    if (rep) {
        std::memset(rep, 0, sizeof *rep);
        rep->solves = h.solves;
        rep->inner_iterations = uint32_t(h.inner);
        rep->outer_sweeps = uint32_t(h.sweeps);
        rep->active_rows = h.active;
        rep->last_relative_residual = h.last_rel;
        rep->last_angle = h.last_angle;
        rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rep->fracture = h.frac;
    }
    // Later stage calls see the substep length and the solve parameters.
    h.params.fixed_dt = prm->fixed_dt / prm->substeps;
    h.phase = P_IDLE;
    return status;
}
