// Device-resident structured scenes: rotation fit, bond assembly and the
// mixed-precision PCG solve on an nx*ny*nz REV box.
//
// HBM layout (all structure-of-arrays planes, plane q of a Q-wide field at
// base + q*n, voxel v = x + nx*(y + ny*z)):
//   state      positions/predicted/previous/velocity/drift 3 x f64,
//              rotations 4 x f64, material constants f32, bond flags u8
//   operator   per +axis bond a: k (symmetric 3x3, 6 values) in f64 and f32,
//              bond force 3 x f64; per voxel: projected diagonal D (6 values)
//              in f64 and f32, its inverse in f64 and f32, rhs 3 x f64, free
//              mask u8
//   solve      x (FP64 solution, double-buffered), d, r, z, p (x2), Ap in the
//              inner precision
// Each bond block is stored once, on its -end voxel; the SpMV of a row reads
// its own three +axis blocks and the -axis neighbours' blocks (the matrix is
// symmetric, A_ji = A_ij^T = A_ij). The free mask is applied to the row
// output; every PCG vector is zero on fixed DOFs, so no column projection
// is needed.
//
// The algorithm restates src/solver.rs:27-110 (frames, compliance,
// stiffness, thermal rest), :327-417 (rotation fit), :1325-1345 (constraint)
// and :1552-1636 (assembly and projection); the PCG semantics follow
// src/solver.rs:1833-1911 with block-Jacobi, run as FP32 inner solves inside
// an FP64 true-residual refinement loop.
#include "vx_common.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace vx {
namespace {

constexpr double kCell = 0.01;  // src/scene.rs:10
constexpr uint32_t kInnerChunk = 16;

// ---------------------------------------------------------------------------
// FP64 small linear algebra (columns c0, c1, c2 like the reference's Mat3).
// ---------------------------------------------------------------------------
struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot3(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ d3 cross3(d3 a, d3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ d3 normalize_or_zero(d3 a) {
    double rcp = 1.0 / sqrt(dot3(a, a));
    if (isfinite(rcp) && rcp > 0.0) return a * rcp;
    return {0.0, 0.0, 0.0};
}
struct m3 {
    d3 c0, c1, c2;
};
__device__ __forceinline__ d3 col(const m3 &m, int i) { return i == 0 ? m.c0 : (i == 1 ? m.c1 : m.c2); }
__device__ __forceinline__ d3 mul(const m3 &m, d3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }
__device__ __forceinline__ d3 mul_t(const m3 &m, d3 v) { return {dot3(m.c0, v), dot3(m.c1, v), dot3(m.c2, v)}; }
__device__ __forceinline__ double det3(const m3 &m) { return dot3(m.c2, cross3(m.c0, m.c1)); }
__device__ __forceinline__ m3 transpose3(const m3 &m) {
    return {{m.c0.x, m.c1.x, m.c2.x}, {m.c0.y, m.c1.y, m.c2.y}, {m.c0.z, m.c1.z, m.c2.z}};
}
// Inverse by cross products of column pairs (rows of the inverse).
__device__ __forceinline__ m3 inverse3m(const m3 &m) {
    d3 t0 = cross3(m.c1, m.c2), t1 = cross3(m.c2, m.c0), t2 = cross3(m.c0, m.c1);
    double inv = 1.0 / dot3(m.c2, t2);
    return transpose3(m3{t0 * inv, t1 * inv, t2 * inv});
}

struct q4 {
    double x, y, z, w;
};
__device__ __forceinline__ double qdot(q4 a, q4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
__device__ __forceinline__ q4 qnormalize(q4 q) {
    double r = 1.0 / sqrt(qdot(q, q));
    return {q.x * r, q.y * r, q.z * r, q.w * r};
}
__device__ __forceinline__ d3 qrotate(q4 q, d3 v) {
    d3 b{q.x, q.y, q.z};
    return v * (q.w * q.w - dot3(b, b)) + b * (dot3(v, b) * 2.0) + cross3(b, v) * (q.w * 2.0);
}
__device__ __forceinline__ m3 mat_from_quat(q4 q) {
    double x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    double xx = q.x * x2, xy = q.x * y2, xz = q.x * z2, yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    double wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    return {{1.0 - (yy + zz), xy + wz, xz - wy}, {xy - wz, 1.0 - (xx + zz), yz + wx}, {xz + wy, yz - wx, 1.0 - (xx + yy)}};
}
// Shepperd's method: pick the largest of w^2, x^2, y^2, z^2 for the root.
__device__ __forceinline__ q4 quat_from_mat(const m3 &m) {
    double m00 = m.c0.x, m01 = m.c0.y, m02 = m.c0.z, m10 = m.c1.x, m11 = m.c1.y, m12 = m.c1.z;
    double m20 = m.c2.x, m21 = m.c2.y, m22 = m.c2.z;
    if (m22 <= 0.0) {
        double dif10 = m11 - m00, omm22 = 1.0 - m22;
        if (dif10 <= 0.0) {
            double f = omm22 - dif10, s = 0.5 / sqrt(f);
            return {f * s, (m01 + m10) * s, (m02 + m20) * s, (m12 - m21) * s};
        }
        double f = omm22 + dif10, s = 0.5 / sqrt(f);
        return {(m01 + m10) * s, f * s, (m12 + m21) * s, (m20 - m02) * s};
    }
    double sum10 = m11 + m00, opm22 = 1.0 + m22;
    if (sum10 <= 0.0) {
        double f = opm22 - sum10, s = 0.5 / sqrt(f);
        return {(m02 + m20) * s, (m12 + m21) * s, f * s, (m01 - m10) * s};
    }
    double f = opm22 + sum10, s = 0.5 / sqrt(f);
    return {(m12 - m21) * s, (m20 - m02) * s, (m01 - m10) * s, f * s};
}
// Half-way slerp with the shortest-arc flip; normalized lerp above 0.9995.
__device__ __forceinline__ q4 slerp_half(q4 a, q4 b) {
    double d = qdot(a, b);
    if (d < 0.0) {
        b = {-b.x, -b.y, -b.z, -b.w};
        d = -d;
    }
    if (d > 0.9995) return qnormalize({a.x + (b.x - a.x) * 0.5, a.y + (b.y - a.y) * 0.5, a.z + (b.z - a.z) * 0.5, a.w + (b.w - a.w) * 0.5});
    double th = acos(d);
    double s1 = sin(th * 0.5), s2 = sin(th * 0.5), inv = 1.0 / sin(th);
    return {(a.x * s1 + b.x * s2) * inv, (a.y * s1 + b.y * s2) * inv, (a.z * s1 + b.z * s2) * inv, (a.w * s1 + b.w * s2) * inv};
}
// basis(axis, q): rotated e_axis, e_axis+1, e_axis+2 (src/solver.rs:27-34).
__device__ __forceinline__ m3 basis(int axis, q4 q) {
    const d3 e[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    return {qrotate(q, e[axis]), qrotate(q, e[(axis + 1) % 3]), qrotate(q, e[(axis + 2) % 3])};
}

// Symmetric 3x3 stored as (xx, xy, xz, yy, yz, zz).
struct s6 {
    double a[6];
};

// ---------------------------------------------------------------------------
// Scene storage
// ---------------------------------------------------------------------------
struct Geo {
    uint32_t nx, ny, nz, n;
    __device__ __forceinline__ void coords(uint32_t v, uint32_t &x, uint32_t &y, uint32_t &z) const {
        x = v % nx;
        uint32_t t = v / nx;
        y = t % ny;
        z = t / ny;
    }
    __device__ __forceinline__ uint32_t stride(int a) const { return a == 0 ? 1u : (a == 1 ? nx : nx * ny); }
};

// Device-resident solve scalars. Each field is written by workgroup 0 of one
// kernel and read only by later kernels.
struct GridState {
    double rho[2];         // r.z after k inner updates, slot k & 1 (grid_spmv)
    double rr0;            // r.r at the start of the inner solve (grid_spmv, k = 0)
    double inner_tol;      // relative inner tolerance
    double rr_last;        // r.r seen by the last check
    uint64_t angle_bits;   // max rotation change as ordered bits of a non-negative double
    uint32_t k_spmv;       // inner updates completed (grid_update, grid_residual)
    uint32_t k_update;     // iteration of the current direction (grid_spmv)
    uint32_t iters, done, max_iter, pad;
};
enum : uint32_t { kRun = 0, kConv = 1, kNonFin = 2, kCap = 3 };
enum : int { kPap = 0, kRz = 1, kRr = 2, kRr64 = 3, kBb = 4 };  // partial slots

struct DevScene {
    Geo g;
    // state
    double *pos, *pred, *prev, *vel, *drift, *rot;
    float *mass, *young, *young_perp, *shear, *solid, *fiber, *aniso, *dtemp, *a_par, *a_perp;
    uint8_t *kinematic, *bonds, *tmask;  // bonds: bit a = +axis bond present; tmask: bit a = axis targeted
    double *targets;
    float *loads;
    float *damage, *kscale, *offset, *plastic;  // 3*n, 3*n, 9*n ([a][c][v]), 9*n
    // operator
    double *k64, *force;  // 18*n ([a][q][v]), 9*n ([a][c][v])
    float *k32;           // 18*n
    double *d64, *dinv64, *rhs;  // 6*n, 6*n, 3*n
    float *d32, *dinv32;         // 6*n, 6*n
    uint8_t *mask;
    // solve
    double *xa, *xb;             // FP64 solution, ping-pong
    float *d, *r, *z, *p0, *p1, *ap;  // FP32 inner
    double *dd, *rd, *zd, *p0d, *p1d, *apd;  // FP64 inner (plain FP64 mode)
    double *partials;
    GridState *st;
};

}  // namespace
}  // namespace vx

struct vx_scene {
    vx::DevScene s;
    uint32_t active = 0;
    uint32_t per_xcd = 0, ntiles = 0, grid = 0;
    void *alloc = nullptr;
    size_t alloc_bytes = 0;
    bool x_in_a = true;           // which ping-pong buffer holds the current solution
    bool has_solution = false;
    hipGraphExec_t inner32 = nullptr, inner64 = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    vx::GridState *pinned = nullptr;  // 3 status slots
    double *host_partials = nullptr;  // pinned, 2 * grid
};

namespace vx {
namespace {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ d3 ld3(const double *p, size_t n, uint32_t v) { return {p[v], p[n + v], p[2 * n + v]}; }
__device__ __forceinline__ void st3(double *p, size_t n, uint32_t v, d3 a) {
    p[v] = a.x;
    p[n + v] = a.y;
    p[2 * n + v] = a.z;
}
__device__ __forceinline__ q4 ldq(const double *p, size_t n, uint32_t v) { return {p[v], p[n + v], p[2 * n + v], p[3 * n + v]}; }

// Predict one substep (src/solver.rs:494-523, 542).
__global__ __launch_bounds__(kBlock) void grid_predict(DevScene S, double h, double gx, double gy, double gz) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    const size_t n = S.g.n;
    if (v >= n) return;
    d3 p = ld3(S.pos, n, v);
    st3(S.prev, n, v, p);
    const uint8_t tm = S.tmask[v];
    if (!S.kinematic[v]) {
        d3 vel = ld3(S.vel, n, v) + mk(gx, gy, gz) * h;
        if (tm & 1) vel.x = 0.0;
        if (tm & 2) vel.y = 0.0;
        if (tm & 4) vel.z = 0.0;
        st3(S.vel, n, v, vel);
        st3(S.drift, n, v, vel);
        p = p + vel * h;
    } else {
        st3(S.drift, n, v, mk(0, 0, 0));
    }
    if (tm & 1) p.x = S.targets[v];
    if (tm & 2) p.y = S.targets[n + v];
    if (tm & 4) p.z = S.targets[2 * n + v];
    st3(S.pos, n, v, p);
    st3(S.pred, n, v, p);
}

// Rotation fit (src/solver.rs:327-417) in FP64.
__global__ __launch_bounds__(kBlock) void grid_fit(DevScene S) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    const size_t n = S.g.n;
    double angle = 0.0;
    if (v < n && !S.kinematic[v]) {
        uint32_t x, y, z;
        S.g.coords(v, x, y, z);
        const uint32_t c[3] = {x, y, z};
        const q4 prev = ldq(S.rot, n, v);
        const m3 old = mat_from_quat(prev);
        const d3 pv = ld3(S.pos, n, v);
        m3 f;
        d3 *cols[3] = {&f.c0, &f.c1, &f.c2};
        for (int a = 0; a < 3; ++a) {
            d3 sum{0, 0, 0};
            double cnt = 0.0;
            const uint32_t s = S.g.stride(a);
            if (c[a] > 0 && (S.bonds[v - s] >> a & 1)) {
                sum = sum + (pv - ld3(S.pos, n, v - s)) * (1.0 / kCell);
                cnt += 1.0;
            }
            if (S.bonds[v] >> a & 1) {
                sum = sum + (ld3(S.pos, n, v + s) - pv) * (1.0 / kCell);
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
                S.rot[v] = q.x;
                S.rot[n + v] = q.y;
                S.rot[2 * n + v] = q.z;
                S.rot[3 * n + v] = q.w;
            }
            angle = 2.0 * acos(fmin(fmax(fabs(qdot(prev, q)), 0.0), 1.0));
        }
    }
    angle = wave_max(angle);
    if ((threadIdx.x & 63) == 0 && angle > 0.0)
        atomicMax(reinterpret_cast<unsigned long long *>(&S.st->angle_bits), (unsigned long long)__double_as_longlong(angle));
}

// tensor_compliance (src/solver.rs:37-63) of voxel v in the bond frame.
__device__ __forceinline__ bool compliance(const DevScene &S, uint32_t v, const m3 &F, q4 rv, double (&C)[6]) {
    const size_t n = S.g.n;
    double solid = fmin(fmax(double(S.solid[v]), 1e-4), 1.0);
    double E = S.young[v], Ep = S.young_perp[v], G = S.shear[v], an = S.aniso[v];
    d3 fib = mul_t(F, qrotate(rv, mk(S.fiber[v], S.fiber[n + v], S.fiber[2 * n + v])));
    double tr = E + an * (Ep - E);
    double nrm = tr + (E - tr) * fib.x * fib.x;
    d3 dg = mk(nrm, G, G) * solid;
    if (!(isfinite(dg.x) && isfinite(dg.y) && isfinite(dg.z)) || fmin(fmin(dg.x, dg.y), dg.z) <= 0.0) return false;
    d3 inv{1.0 / dg.x, 1.0 / dg.y, 1.0 / dg.z};
    d3 sf{fib.x * inv.x, fib.y * inv.y, fib.z * inv.z};
    double alpha = G * an * 0.2;
    double c = alpha / (1.0 + alpha * dot3(fib, sf));
    C[0] = inv.x - sf.x * sf.x * c;
    C[1] = -sf.x * sf.y * c;
    C[2] = -sf.x * sf.z * c;
    C[3] = inv.y - sf.y * sf.y * c;
    C[4] = -sf.y * sf.z * c;
    C[5] = inv.z - sf.z * sf.z * c;
    return true;
}

__device__ __forceinline__ d3 endpoint_expansion(const DevScene &S, uint32_t v, q4 rv, d3 dir) {
    const size_t n = S.g.n;
    d3 fib = normalize_or_zero(qrotate(rv, mk(S.fiber[v], S.fiber[n + v], S.fiber[2 * n + v])));
    double ap = S.a_par[v], aq = S.a_perp[v];
    d3 ad = dir * aq + fib * ((ap - aq) * dot3(fib, dir));
    return ad * double(S.dtemp[v]);
}

// Per +axis bond: rotated stiffness k and force k (p_j - p_i - rest)
// (src/solver.rs:1325-1345, 1595-1619).
__global__ __launch_bounds__(kBlock) void grid_bonds(DevScene S) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    const size_t n = S.g.n;
    if (v >= n) return;
    const q4 qi = ldq(S.rot, n, v);
    const d3 pi = ld3(S.pos, n, v);
    const uint8_t mi = S.mask[v];
    for (int a = 0; a < 3; ++a) {
        double k[6] = {0, 0, 0, 0, 0, 0};
        d3 force{0, 0, 0};
        const uint32_t j = v + S.g.stride(a);
        if ((S.bonds[v] >> a & 1) && (mi | S.mask[j])) {
            const q4 qj = ldq(S.rot, n, j);
            const m3 F = basis(a, slerp_half(qi, qj));
            double Ca[6], Cb[6];
            if (compliance(S, v, F, qi, Ca) && compliance(S, j, F, qj, Cb)) {
                // K = (Ca + Cb)^-1 * 2*AREA/CELL, symmetric by cofactors.
                double s0 = Ca[0] + Cb[0], s1 = Ca[1] + Cb[1], s2 = Ca[2] + Cb[2];
                double s3 = Ca[3] + Cb[3], s4 = Ca[4] + Cb[4], s5 = Ca[5] + Cb[5];
                double c00 = s3 * s5 - s4 * s4, c01 = s2 * s4 - s1 * s5, c02 = s1 * s4 - s2 * s3;
                double c11 = s0 * s5 - s2 * s2, c12 = s1 * s2 - s0 * s4, c22 = s0 * s3 - s1 * s1;
                double det = s0 * c00 + s1 * c01 + s2 * c02;
                const size_t b = size_t(a) * n + v;
                double scale = (2.0 * kCell * kCell / kCell) / det * double(S.kscale[b]) *
                               fmax(1.0 - double(S.damage[b]), 0.001);
                double K[3][3] = {{c00 * scale, c01 * scale, c02 * scale},
                                  {c01 * scale, c11 * scale, c12 * scale},
                                  {c02 * scale, c12 * scale, c22 * scale}};
                // k = F K F^T
                double Fm[3][3] = {{F.c0.x, F.c1.x, F.c2.x}, {F.c0.y, F.c1.y, F.c2.y}, {F.c0.z, F.c1.z, F.c2.z}};
                double T[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) T[r][c] = Fm[r][0] * K[0][c] + Fm[r][1] * K[1][c] + Fm[r][2] * K[2][c];
                int q = 0;
                for (int r = 0; r < 3; ++r)
                    for (int c = r; c < 3; ++c) k[q++] = T[r][0] * Fm[c][0] + T[r][1] * Fm[c][1] + T[r][2] * Fm[c][2];
            }
            // rest = F (thermal rest + formation offset + plastic * CELL)
            d3 dir = normalize_or_zero(F.c0);
            d3 therm = mul_t(F, dir * kCell + (endpoint_expansion(S, v, qi, dir) + endpoint_expansion(S, j, qj, dir)) * (0.5 * kCell));
            const size_t o = size_t(a) * 3 * n;
            d3 off = mk(S.offset[o + v], S.offset[o + n + v], S.offset[o + 2 * n + v]);
            d3 pl = mk(S.plastic[o + v], S.plastic[o + n + v], S.plastic[o + 2 * n + v]);
            d3 rest = mul(F, therm + off + pl * kCell);
            d3 e = ld3(S.pos, n, j) - pi - rest;
            force = {k[0] * e.x + k[1] * e.y + k[2] * e.z, k[1] * e.x + k[3] * e.y + k[4] * e.z,
                     k[2] * e.x + k[4] * e.y + k[5] * e.z};
        }
        for (int q = 0; q < 6; ++q) {
            S.k64[(size_t(a) * 6 + q) * n + v] = k[q];
            S.k32[(size_t(a) * 6 + q) * n + v] = float(k[q]);
        }
        S.force[(size_t(a) * 3 + 0) * n + v] = force.x;
        S.force[(size_t(a) * 3 + 1) * n + v] = force.y;
        S.force[(size_t(a) * 3 + 2) * n + v] = force.z;
    }
}

__device__ __forceinline__ void sym_inverse(const double (&a)[6], double (&o)[6]) {
    double c00 = a[3] * a[5] - a[4] * a[4], c01 = a[2] * a[4] - a[1] * a[5], c02 = a[1] * a[4] - a[2] * a[3];
    double c11 = a[0] * a[5] - a[2] * a[2], c12 = a[1] * a[2] - a[0] * a[4], c22 = a[0] * a[3] - a[1] * a[1];
    double inv = 1.0 / (a[0] * c00 + a[1] * c01 + a[2] * c02);
    o[0] = c00 * inv;
    o[1] = c01 * inv;
    o[2] = c02 * inv;
    o[3] = c11 * inv;
    o[4] = c12 * inv;
    o[5] = c22 * inv;
}

// Per voxel: D = m/h^2 I + sum k, rhs = (pred - p) m/h^2 + load +- forces,
// projection D <- P D P + I - P, rhs <- P rhs (src/solver.rs:1576-1636).
__global__ __launch_bounds__(kBlock) void grid_rows(DevScene S, double h) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    const size_t n = S.g.n;
    if (v >= n) return;
    uint32_t x, y, z;
    S.g.coords(v, x, y, z);
    const uint32_t c[3] = {x, y, z};
    const double m = double(S.mass[v]) / (h * h);
    double D[6] = {m, 0, 0, m, 0, m};
    d3 rhs = (ld3(S.pred, n, v) - ld3(S.pos, n, v)) * m;
    if (!S.kinematic[v]) rhs = rhs + mk(S.loads[v], S.loads[n + v], S.loads[2 * n + v]);
    for (int a = 0; a < 3; ++a) {
        const uint32_t s = S.g.stride(a);
        if (c[a] > 0) {  // bond (v - s -> v): this voxel is its +end
            const uint32_t u = v - s;
            for (int q = 0; q < 6; ++q) D[q] += S.k64[(size_t(a) * 6 + q) * n + u];
            rhs = rhs - mk(S.force[(size_t(a) * 3) * n + u], S.force[(size_t(a) * 3 + 1) * n + u], S.force[(size_t(a) * 3 + 2) * n + u]);
        }
        for (int q = 0; q < 6; ++q) D[q] += S.k64[(size_t(a) * 6 + q) * n + v];
        rhs = rhs + mk(S.force[(size_t(a) * 3) * n + v], S.force[(size_t(a) * 3 + 1) * n + v], S.force[(size_t(a) * 3 + 2) * n + v]);
    }
    const uint8_t mk_ = S.mask[v];
    const double f[3] = {double(mk_ & 1), double(mk_ >> 1 & 1), double(mk_ >> 2 & 1)};
    // P D P + I - P on the symmetric storage
    const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double Dp[6];
    for (int r = 0; r < 3; ++r)
        for (int cc = r; cc < 3; ++cc) Dp[idx[r][cc]] = f[r] * D[idx[r][cc]] * f[cc] + (r == cc ? 1.0 - f[r] : 0.0);
    double Di[6];
    sym_inverse(Dp, Di);
    for (int q = 0; q < 6; ++q) {
        S.d64[q * n + v] = Dp[q];
        S.d32[q * n + v] = float(Dp[q]);
        S.dinv64[q * n + v] = Di[q];
        S.dinv32[q * n + v] = float(Di[q]);
    }
    st3(S.rhs, n, v, mk(rhs.x * f[0], rhs.y * f[1], rhs.z * f[2]));
}

// ---- PCG kernels on the structured operator --------------------------------
template <class S> __device__ __forceinline__ S fmaS(S a, S b, S c);
template <> __device__ __forceinline__ float fmaS<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fmaS<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class S>
__device__ __forceinline__ void symv(const S *__restrict__ m, size_t n, uint32_t v, const S (&x)[3], S (&y)[3]) {
    const S a0 = m[v], a1 = m[n + v], a2 = m[2 * n + v], a3 = m[3 * n + v], a4 = m[4 * n + v], a5 = m[5 * n + v];
    y[0] = fmaS<S>(a2, x[2], fmaS<S>(a1, x[1], a0 * x[0]));
    y[1] = fmaS<S>(a4, x[2], fmaS<S>(a3, x[1], a1 * x[0]));
    y[2] = fmaS<S>(a5, x[2], fmaS<S>(a4, x[1], a2 * x[0]));
}
template <class S>
__device__ __forceinline__ void symv_acc(const S *__restrict__ m, size_t n, uint32_t v, const S (&x)[3], S (&y)[3]) {
    const S a0 = m[v], a1 = m[n + v], a2 = m[2 * n + v], a3 = m[3 * n + v], a4 = m[4 * n + v], a5 = m[5 * n + v];
    y[0] = fmaS<S>(a2, x[2], fmaS<S>(a1, x[1], fmaS<S>(a0, x[0], y[0])));
    y[1] = fmaS<S>(a4, x[2], fmaS<S>(a3, x[1], fmaS<S>(a1, x[0], y[1])));
    y[2] = fmaS<S>(a5, x[2], fmaS<S>(a4, x[1], fmaS<S>(a2, x[0], y[2])));
}

template <class S> struct Inner;
template <> struct Inner<float> {
    static __device__ __forceinline__ const float *k(const DevScene &s) { return s.k32; }
    static __device__ __forceinline__ const float *d(const DevScene &s) { return s.d32; }
    static __device__ __forceinline__ const float *dinv(const DevScene &s) { return s.dinv32; }
};
template <> struct Inner<double> {
    static __device__ __forceinline__ const double *k(const DevScene &s) { return s.k64; }
    static __device__ __forceinline__ const double *d(const DevScene &s) { return s.d64; }
    static __device__ __forceinline__ const double *dinv(const DevScene &s) { return s.dinv64; }
};

// Inner convergence check of the previous update (every iteration:
// r.r <= tol^2 r0.r0), beta, then y = A p on the structured 7-point
// 3x3-block operator with p = z + beta p_old formed on load for this row and
// its neighbours; partials of p.Ap.
template <class S>
__global__ __launch_bounds__(kBlock) void grid_spmv(DevScene G, uint32_t per_xcd, uint32_t ntiles, const S *__restrict__ z,
                                                    const S *__restrict__ p_old, S *__restrict__ p_new, S *__restrict__ ap) {
    GridState *st = G.st;
    __shared__ double lds[2][kWaves];
    // Scalars first: the state words and the previous update's partials are
    // loaded together, then summed across the workgroup.
    const uint32_t done_in = st->done;
    const uint32_t k = st->k_spmv;
    const uint32_t max_iter = st->max_iter;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const double rr0_s = st->rr0, tol = st->inner_tol;
    const uint32_t nb = gridDim.x;
    double sums[2];
    const int in_slots[2] = {kRz, kRr};
    sum_partials<2>(sums, G.partials, in_slots, nb, nb, lds);
    if (done_in) return;
    const double rho = sums[0], rr = sums[1];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double rr0 = k == 0 ? rr : rr0_s;
    uint32_t done = kRun;
    if (!isfinite(rr))
        done = kNonFin;
    else if (!(rr > tol * tol * rr0))
        done = kConv;
    else if (k >= max_iter)
        done = kCap;
    if (lead) {
        st->rr_last = rr;
        if (k == 0) st->rr0 = rr;
    }
    if (done) {
        if (lead) {
            st->done = done;
            st->iters = k;
        }
        return;
    }
    const double rho_prev = k == 0 ? 0.0 : rho_s[(k - 1) & 1];
    const S beta = (k == 0 || !(rho_prev > 1e-300)) ? S(0) : S(rho / rho_prev);
    if (lead) {
        st->rho[k & 1] = rho;
        st->k_update = k;
    }
    const size_t n = G.g.n;
    const uint32_t tile = xcd_tile(blockIdx.x, per_xcd);
    const uint32_t v = tile * kBlock + threadIdx.x;
    double part[1] = {0.0};
    if (tile < ntiles && v < n) {
        const S *__restrict__ K = Inner<S>::k(G);
        uint32_t x, y, zc;
        G.g.coords(v, x, y, zc);
        const uint32_t c[3] = {x, y, zc};
        const uint32_t ext[3] = {G.g.nx, G.g.ny, G.g.nz};
        // Load everything first (zeros where a neighbour is absent) so the
        // loads of all six neighbours are in flight together; compute after.
        S zz[7][3], pp[7][3], Dv[6], Kp[3][6], Km[3][6];
        const uint8_t m = G.mask[v];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            zz[0][q] = z[q * n + v];
            pp[0][q] = p_old[q * n + v];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) Dv[q] = Inner<S>::d(G)[q * n + v];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = G.g.stride(a);
            const S *Ka = K + size_t(a) * 6 * n;
            const bool lo = c[a] > 0, hi = c[a] + 1 < ext[a];
            const uint32_t um = lo ? v - s : v, up = hi ? v + s : v;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                zz[1 + 2 * a][q] = lo ? z[q * n + um] : S(0);
                pp[1 + 2 * a][q] = lo ? p_old[q * n + um] : S(0);
                zz[2 + 2 * a][q] = hi ? z[q * n + up] : S(0);
                pp[2 + 2 * a][q] = hi ? p_old[q * n + up] : S(0);
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                Km[a][q] = lo ? Ka[q * n + um] : S(0);
                Kp[a][q] = Ka[q * n + v];  // zero when the +axis bond is absent
            }
        }
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
        p_new[v] = p[0][0];
        p_new[n + v] = p[0][1];
        p_new[2 * n + v] = p[0][2];
        ap[v] = y0;
        ap[n + v] = y1;
        ap[2 * n + v] = y2;
        part[0] = double(p[0][0]) * double(y0);
        part[0] = __builtin_fma(double(p[0][1]), double(y1), part[0]);
        part[0] = __builtin_fma(double(p[0][2]), double(y2), part[0]);
    }
    const int out_slots[1] = {kPap};
    store_partials<1>(part, G.partials, out_slots, nb, lds);
}

// alpha = rho / p.Ap; d += alpha p (d = alpha p on the first iteration);
// r -= alpha Ap; z = Dinv r; partials of r.z and r.r.
template <class S>
__global__ __launch_bounds__(kBlock) void grid_update(DevScene G, uint32_t per_xcd, uint32_t ntiles, S *__restrict__ d,
                                                      S *__restrict__ r, S *__restrict__ z, const S *__restrict__ p,
                                                      const S *__restrict__ ap) {
    GridState *st = G.st;
    __shared__ double lds[2][kWaves];
    const uint32_t done_in = st->done;
    const uint32_t k = st->k_update;
    const double rho_s[2] = {st->rho[0], st->rho[1]};
    const uint32_t nb = gridDim.x;
    double pap[1];
    const int in_slots[1] = {kPap};
    sum_partials<1>(pap, G.partials, in_slots, nb, nb, lds);
    if (done_in) return;
    const S alpha = pap[0] > 1e-300 ? S(rho_s[k & 1] / pap[0]) : S(0);
    const bool first = k == 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->k_spmv = k + 1;
        st->iters = k + 1;
    }
    const size_t n = G.g.n;
    const uint32_t tile = xcd_tile(blockIdx.x, per_xcd);
    const uint32_t v = tile * kBlock + threadIdx.x;
    double part[2] = {0.0, 0.0};
    if (tile < ntiles && v < n) {
        S pv[3], av[3], rv[3], dv[3], Di[6];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            pv[q] = p[q * n + v];
            av[q] = ap[q * n + v];
            rv[q] = r[q * n + v];
            dv[q] = first ? S(0) : d[q * n + v];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) Di[q] = Inner<S>::dinv(G)[q * n + v];
        S ri[3], zi[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            d[q * n + v] = fmaS<S>(alpha, pv[q], dv[q]);
            ri[q] = fmaS<S>(-alpha, av[q], rv[q]);
            r[q * n + v] = ri[q];
        }
        zi[0] = fmaS<S>(Di[2], ri[2], fmaS<S>(Di[1], ri[1], Di[0] * ri[0]));
        zi[1] = fmaS<S>(Di[4], ri[2], fmaS<S>(Di[3], ri[1], Di[1] * ri[0]));
        zi[2] = fmaS<S>(Di[5], ri[2], fmaS<S>(Di[4], ri[1], Di[2] * ri[0]));
#pragma unroll
        for (int q = 0; q < 3; ++q) z[q * n + v] = zi[q];
        part[0] = __builtin_fma(double(ri[2]), double(zi[2]), __builtin_fma(double(ri[1]), double(zi[1]), double(ri[0]) * double(zi[0])));
        part[1] = __builtin_fma(double(ri[2]), double(ri[2]), __builtin_fma(double(ri[1]), double(ri[1]), double(ri[0]) * double(ri[0])));
    }
    const int out_slots[2] = {kRz, kRr};
    store_partials<2>(part, G.partials, out_slots, nb, lds);
}

// Outer FP64 refinement step: x_new = x_old + d (own and neighbour rows),
// r64 = b - A64 x_new; the inner system starts from r = r64 in the inner
// precision, z = Dinv r, p_old = 0. Partials of ||r64||^2, ||b||^2, r.z, r.r.
template <class S>
__global__ __launch_bounds__(kBlock) void grid_residual(DevScene G, uint32_t per_xcd, uint32_t ntiles, const double *__restrict__ x_old,
                                                        double *__restrict__ x_new, const S *__restrict__ d, int use_d,
                                                        S *__restrict__ r, S *__restrict__ z, S *__restrict__ p_old) {
    GridState *st = G.st;
    __shared__ double lds[4][kWaves];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->k_spmv = 0;
        st->done = kRun;
    }
    const size_t n = G.g.n;
    const uint32_t tile = xcd_tile(blockIdx.x, per_xcd);
    const uint32_t v = tile * kBlock + threadIdx.x;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    if (tile < ntiles && v < n) {
        uint32_t x, y, zc;
        G.g.coords(v, x, y, zc);
        const uint32_t c[3] = {x, y, zc};
        const uint32_t ext[3] = {G.g.nx, G.g.ny, G.g.nz};
        auto xval = [&](uint32_t u, double (&o)[3]) {
#pragma unroll
            for (int q = 0; q < 3; ++q) o[q] = x_old[q * n + u] + (use_d ? double(d[q * n + u]) : 0.0);
        };
        double xv[3];
        xval(v, xv);
        double yv[3];
        symv<double>(G.d64, n, v, xv, yv);
        double acc[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t s = G.g.stride(a);
            if (c[a] > 0) {
                double xu[3];
                xval(v - s, xu);
                symv_acc<double>(G.k64 + size_t(a) * 6 * n, n, v - s, xu, acc);
            }
            if (c[a] + 1 < ext[a]) {
                double xu[3];
                xval(v + s, xu);
                symv_acc<double>(G.k64 + size_t(a) * 6 * n, n, v, xu, acc);
            }
        }
        const uint8_t m = G.mask[v];
        const double f[3] = {double(m & 1), double(m >> 1 & 1), double(m >> 2 & 1)};
        S rs[3], zs[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double b = G.rhs[q * n + v];
            const double r64 = f[q] * (b - (yv[q] - acc[q]));
            x_new[q * n + v] = xv[q];
            rs[q] = S(r64);
            r[q * n + v] = rs[q];
            p_old[q * n + v] = S(0);
            part[0] = __builtin_fma(r64, r64, part[0]);
            part[1] = __builtin_fma(b, b, part[1]);
        }
        symv<S>(Inner<S>::dinv(G), n, v, rs, zs);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            z[q * n + v] = zs[q];
            part[2] = __builtin_fma(double(rs[q]), double(zs[q]), part[2]);
            part[3] = __builtin_fma(double(rs[q]), double(rs[q]), part[3]);
        }
    }
    const int out_slots[4] = {kRr64, kBb, kRz, kRr};
    store_partials<4>(part, G.partials, out_slots, gridDim.x, lds);
}

// positions += x on the active rows (src/solver.rs:1921-1923).
__global__ __launch_bounds__(kBlock) void grid_apply(DevScene S, const double *__restrict__ x) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    const size_t n = S.g.n;
    if (v >= n || S.mask[v] == 0) return;
    for (int q = 0; q < 3; ++q) S.pos[q * n + v] += x[q * n + v];
}

// Velocities after a substep (src/solver.rs:668-683), no contact.
__global__ __launch_bounds__(kBlock) void grid_finish(DevScene S, double h) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    const size_t n = S.g.n;
    if (v >= n || S.kinematic[v]) return;
    const uint8_t tm = S.tmask[v];
    for (int q = 0; q < 3; ++q) {
        const double p = S.pos[q * n + v];
        if (tm >> q & 1)
            S.vel[q * n + v] = (p - S.prev[q * n + v]) / h;
        else
            S.vel[q * n + v] = S.drift[q * n + v] + (p - S.pred[q * n + v]) / h;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int32_t check_scene(vx_scene *sc) {
    if (!sc) return fail(VX_ERR_INVALID_ARGUMENT, "null scene");
    return VX_OK;
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

int32_t read_state(vx_scene *sc, GridState *out) {
    Context &ctx = context();
    VX_HIP(hipMemcpyAsync(&sc->pinned[2], sc->s.st, sizeof(GridState), hipMemcpyDeviceToHost, ctx.stream));
    VX_HIP(hipStreamSynchronize(ctx.stream));
    *out = sc->pinned[2];
    return VX_OK;
}

template <class S>
int32_t capture_inner(vx_scene *sc, hipGraphExec_t *exec) {
    Context &ctx = context();
    DevScene &G = sc->s;
    S *d, *r, *z, *p0, *p1, *ap;
    if (sizeof(S) == 4) {
        d = (S *)G.d, r = (S *)G.r, z = (S *)G.z, p0 = (S *)G.p0, p1 = (S *)G.p1, ap = (S *)G.ap;
    } else {
        d = (S *)G.dd, r = (S *)G.rd, z = (S *)G.zd, p0 = (S *)G.p0d, p1 = (S *)G.p1d, ap = (S *)G.apd;
    }
    hipGraph_t graph;
    VX_HIP(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t t = 0; t < kInnerChunk; ++t) {
        S *pold = (t % 2 == 0) ? p0 : p1;
        S *pnew = (t % 2 == 0) ? p1 : p0;
        hipLaunchKernelGGL(grid_spmv<S>, dim3(sc->grid), dim3(kBlock), 0, ctx.stream, G, sc->per_xcd, sc->ntiles, z, pold, pnew, ap);
        hipLaunchKernelGGL(grid_update<S>, dim3(sc->grid), dim3(kBlock), 0, ctx.stream, G, sc->per_xcd, sc->ntiles, d, r, z, pnew, ap);
    }
    VX_HIP(hipStreamEndCapture(ctx.stream, &graph));
    hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(VX_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return VX_OK;
}

// Inner PCG: replay 16-iteration graphs, reading the status one graph behind.
int32_t run_inner(vx_scene *sc, hipGraphExec_t exec, uint32_t max_inner) {
    Context &ctx = context();
    uint32_t launched = 0, chunk = 0;
    // The check after the last update runs in the following grid_spmv, so
    // replay until one spmv past the cap has been issued.
    while (launched <= max_inner) {
        VX_HIP(hipGraphLaunch(exec, ctx.stream));
        VX_HIP(hipMemcpyAsync(&sc->pinned[chunk % 2], sc->s.st, sizeof(GridState), hipMemcpyDeviceToHost, ctx.stream));
        VX_HIP(hipEventRecord(sc->ev[chunk % 2], ctx.stream));
        launched += kInnerChunk;
        if (chunk >= 1) {
            VX_HIP(hipEventSynchronize(sc->ev[(chunk - 1) % 2]));
            if (sc->pinned[(chunk - 1) % 2].done) break;
        }
        chunk += 1;
    }
    return VX_OK;
}

template <class S>
int32_t solve_t(vx_scene *sc, const vx_solve_params *prm, vx_solve_report *rep) {
    Context &ctx = context();
    DevScene &G = sc->s;
    const bool f32 = sizeof(S) == 4;
    hipGraphExec_t &exec = f32 ? sc->inner32 : sc->inner64;
    if (!exec) {
        int32_t c = capture_inner<S>(sc, &exec);
        if (c != VX_OK) return c;
    }
    S *d = f32 ? (S *)G.d : (S *)G.dd;
    S *r = f32 ? (S *)G.r : (S *)G.rd;
    S *z = f32 ? (S *)G.z : (S *)G.zd;
    S *p0 = f32 ? (S *)G.p0 : (S *)G.p0d;
    // Plain FP64: one sweep solved straight to the outer tolerance.
    const double inner_tol = f32 ? prm->inner_tolerance : prm->tolerance * 0.5;
    const uint32_t max_outer = f32 ? prm->max_outer : 2;
    GridState init;
    std::memset(&init, 0, sizeof init);
    init.inner_tol = inner_tol;
    init.max_iter = prm->max_inner;
    sc->pinned[2] = init;
    VX_HIP(hipMemcpyAsync(G.st, &sc->pinned[2], sizeof init, hipMemcpyHostToDevice, ctx.stream));
    const bool warm = prm->warm_start && sc->has_solution;
    if (!warm) VX_HIP(hipMemsetAsync(sc->x_in_a ? G.xa : G.xb, 0, sizeof(double) * 3 * G.g.n, ctx.stream));
    uint32_t inner_total = 0, sweeps = 0;
    double rel = INFINITY;
    int32_t status = VX_OK;
    for (uint32_t s = 0; s <= max_outer; ++s) {
        double *xo = sc->x_in_a ? G.xa : G.xb;
        double *xn = sc->x_in_a ? G.xb : G.xa;
        hipLaunchKernelGGL(grid_residual<S>, dim3(sc->grid), dim3(kBlock), 0, ctx.stream, G, sc->per_xcd, sc->ntiles, xo, xn,
                           (const S *)d, s > 0 ? 1 : 0, r, z, p0);
        VX_HIP(hipGetLastError());
        sc->x_in_a = !sc->x_in_a;
        // ||r64||^2 and ||b||^2 from the residual kernel's partials, summed
        // on the host in a fixed order.
        VX_HIP(hipMemcpyAsync(sc->host_partials, G.partials + size_t(kRr64) * sc->grid, 2 * sizeof(double) * sc->grid,
                              hipMemcpyDeviceToHost, ctx.stream));
        VX_HIP(hipStreamSynchronize(ctx.stream));
        double rr64 = 0.0, bb = 0.0;
        for (uint32_t b = 0; b < sc->grid; ++b) {
            rr64 += sc->host_partials[b];
            bb += sc->host_partials[sc->grid + b];
        }
        sweeps += 1;
        rel = bb > 0.0 ? std::sqrt(rr64 / bb) : 0.0;
        if (rep) rep->rhs_norm = std::sqrt(bb);
        if (!std::isfinite(rr64)) {
            status = fail(VX_ERR_NONFINITE_RESIDUAL, "HIP PCG produced a non-finite residual norm");
            break;
        }
        int32_t c = VX_OK;
        if (rel <= prm->tolerance) break;
        if (s == max_outer) break;
        c = run_inner(sc, exec, prm->max_inner);
        if (c != VX_OK) return c;
        GridState after;
        c = read_state(sc, &after);
        if (c != VX_OK) return c;
        inner_total += after.iters;
        if (after.done == kNonFin) {
            status = fail(VX_ERR_NONFINITE_RESIDUAL, "HIP PCG produced a non-finite residual norm");
            break;
        }
    }
    sc->has_solution = true;
    if (rep) {
        rep->inner_iterations = inner_total;
        rep->outer_sweeps = sweeps;
        rep->active_rows = sc->active;
        rep->relative_residual = rel;
        rep->status = uint32_t(status);
    }
    if (status == VX_OK && !(rel <= prm->tolerance))
        status = fail(VX_ERR_NOT_CONVERGED, "HIP PCG did not converge: inner iterations=%u, sweeps=%u, relative residual=%.3e, tolerance=%.3e",
                      inner_total, sweeps, rel, prm->tolerance);
    if (rep) rep->status = uint32_t(status);
    return status;
}

struct Lock {
    std::lock_guard<std::recursive_mutex> g;
    Lock() : g(context().mutex) {}
};

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" int32_t vx_scene_create(const vx_scene_desc *desc, vx_scene **out) {
    clear_error();
    if (!desc || !out) return fail(VX_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint64_t n64 = uint64_t(desc->nx) * desc->ny * desc->nz;
    if (n64 == 0) return fail(VX_ERR_INVALID_ARGUMENT, "empty scene");
    if (n64 >= (1ull << 31)) return fail(VX_ERR_TOO_LARGE, "scene of %llu REVs exceeds the index range", (unsigned long long)n64);
    if (!desc->positions || !desc->rotations || !desc->velocities || !desc->mass || !desc->kinematic || !desc->young ||
        !desc->young_perp || !desc->shear || !desc->solid || !desc->fibers || !desc->anisotropy || !desc->delta_temp ||
        !desc->alpha_par || !desc->alpha_perp || !desc->neighbor)
        return fail(VX_ERR_INVALID_ARGUMENT, "missing scene array");
    const uint32_t n = uint32_t(n64);
    const uint32_t nx = desc->nx, ny = desc->ny;
    // Bonds must be lattice bonds v -> v + stride_a.
    std::vector<uint8_t> bonds(n, 0), tmask(n, 0), mask(n, 0);
    for (uint32_t a = 0; a < 3; ++a) {
        const uint32_t s = a == 0 ? 1 : (a == 1 ? nx : nx * ny);
        for (uint32_t v = 0; v < n; ++v) {
            uint32_t j = desc->neighbor[size_t(a) * n + v];
            if (j == VX_EMPTY) continue;
            uint32_t c = a == 0 ? v % nx : (a == 1 ? (v / nx) % ny : v / (nx * ny));
            uint32_t e = a == 0 ? nx : (a == 1 ? ny : desc->nz);
            if (c + 1 >= e || j != v + s) return fail(VX_ERR_INVALID_ARGUMENT, "bond %u of axis %u is not a lattice bond", v, a);
            bonds[v] |= uint8_t(1u << a);
        }
    }
    uint32_t active = 0;
    for (uint32_t v = 0; v < n; ++v) {
        uint8_t m = 0;
        for (int a = 0; a < 3; ++a) {
            bool targeted = desc->targets && !std::isnan(desc->targets[3 * size_t(v) + a]);
            if (targeted) tmask[v] |= uint8_t(1u << a);
            if (!desc->kinematic[v] && !targeted) m |= uint8_t(1u << a);
        }
        mask[v] = m;
        active += m ? 1 : 0;
    }
    int32_t c = ensure_init();
    if (c != VX_OK) return c;
    Lock lock;
    Context &ctx = context();
    vx_scene *sc = new vx_scene();
    DevScene &G = sc->s;
    G.g = Geo{desc->nx, desc->ny, desc->nz, n};
    sc->active = active;
    sc->ntiles = blocks_for(n);
    sc->grid = xcd_grid(sc->ntiles, &sc->per_xcd);
    // One allocation, carved into 256-byte aligned planes.
    struct Item {
        void **ptr;
        size_t bytes;
    };
    const size_t N = n;
    std::vector<Item> items = {
        {(void **)&G.pos, 24 * N}, {(void **)&G.pred, 24 * N}, {(void **)&G.prev, 24 * N}, {(void **)&G.vel, 24 * N},
        {(void **)&G.drift, 24 * N}, {(void **)&G.rot, 32 * N}, {(void **)&G.mass, 4 * N}, {(void **)&G.young, 4 * N},
        {(void **)&G.young_perp, 4 * N}, {(void **)&G.shear, 4 * N}, {(void **)&G.solid, 4 * N}, {(void **)&G.fiber, 12 * N},
        {(void **)&G.aniso, 4 * N}, {(void **)&G.dtemp, 4 * N}, {(void **)&G.a_par, 4 * N}, {(void **)&G.a_perp, 4 * N},
        {(void **)&G.kinematic, N}, {(void **)&G.bonds, N}, {(void **)&G.tmask, N}, {(void **)&G.targets, 24 * N},
        {(void **)&G.loads, 12 * N}, {(void **)&G.damage, 12 * N}, {(void **)&G.kscale, 12 * N}, {(void **)&G.offset, 36 * N},
        {(void **)&G.plastic, 36 * N}, {(void **)&G.k64, 144 * N}, {(void **)&G.force, 72 * N}, {(void **)&G.k32, 72 * N},
        {(void **)&G.d64, 48 * N}, {(void **)&G.dinv64, 48 * N}, {(void **)&G.rhs, 24 * N}, {(void **)&G.d32, 24 * N},
        {(void **)&G.dinv32, 24 * N}, {(void **)&G.mask, N}, {(void **)&G.xa, 24 * N}, {(void **)&G.xb, 24 * N},
        {(void **)&G.d, 12 * N}, {(void **)&G.r, 12 * N}, {(void **)&G.z, 12 * N}, {(void **)&G.p0, 12 * N},
        {(void **)&G.p1, 12 * N}, {(void **)&G.ap, 12 * N}, {(void **)&G.dd, 24 * N}, {(void **)&G.rd, 24 * N},
        {(void **)&G.zd, 24 * N}, {(void **)&G.p0d, 24 * N}, {(void **)&G.p1d, 24 * N}, {(void **)&G.apd, 24 * N},
        {(void **)&G.partials, 8 * 8 * size_t(sc->grid)}, {(void **)&G.st, sizeof(GridState)},
    };
    size_t total = 0;
    for (auto &it : items) total += (it.bytes + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&sc->alloc, total);
    if (e != hipSuccess) {
        delete sc;
        return fail(e == hipErrorOutOfMemory ? VX_ERR_TOO_LARGE : VX_ERR_DEVICE, "scene allocation of %zu bytes failed: %s", total,
                    hipGetErrorString(e));
    }
    sc->alloc_bytes = total;
    size_t off = 0;
    for (auto &it : items) {
        *it.ptr = static_cast<char *>(sc->alloc) + off;
        off += (it.bytes + 255) & ~size_t(255);
    }
    auto fail_free = [&](int32_t code) {
        (void)hipFree(sc->alloc);
        delete sc;
        return code;
    };
    // Host staging: AoS inputs -> SoA planes (f64 state).
    std::vector<double> buf(4 * N);
    auto up_d = [&](double *dst, const float *src, int comps) -> int32_t {
        for (int q = 0; q < comps; ++q)
            for (size_t v = 0; v < N; ++v) buf[q * N + v] = src ? double(src[comps * v + q]) : 0.0;
        hipError_t er = hipMemcpy(dst, buf.data(), sizeof(double) * comps * N, hipMemcpyHostToDevice);
        return er == hipSuccess ? VX_OK : fail(VX_ERR_DEVICE, "upload failed: %s", hipGetErrorString(er));
    };
    std::vector<float> fbuf(9 * N);
    auto up_f = [&](float *dst, const float *src, int comps, float dflt) -> int32_t {
        for (int q = 0; q < comps; ++q)
            for (size_t v = 0; v < N; ++v) fbuf[q * N + v] = src ? src[comps * v + q] : dflt;
        hipError_t er = hipMemcpy(dst, fbuf.data(), sizeof(float) * comps * N, hipMemcpyHostToDevice);
        return er == hipSuccess ? VX_OK : fail(VX_ERR_DEVICE, "upload failed: %s", hipGetErrorString(er));
    };
    // Per-bond planes come [axis][voxel] (scalars) or [axis][voxel][3]
    // (vectors); the device wants [axis][component][voxel].
    auto up_bond_vec = [&](float *dst, const float *src) -> int32_t {
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < 3; ++q)
                for (size_t v = 0; v < N; ++v) fbuf[(a * 3 + q) * N + v] = src ? src[(a * N + v) * 3 + q] : 0.0f;
        hipError_t er = hipMemcpy(dst, fbuf.data(), sizeof(float) * 9 * N, hipMemcpyHostToDevice);
        return er == hipSuccess ? VX_OK : fail(VX_ERR_DEVICE, "upload failed: %s", hipGetErrorString(er));
    };
    auto up_bond = [&](float *dst, const float *src, float dflt) -> int32_t {
        for (size_t k = 0; k < 3 * N; ++k) fbuf[k] = src ? src[k] : dflt;
        hipError_t er = hipMemcpy(dst, fbuf.data(), sizeof(float) * 3 * N, hipMemcpyHostToDevice);
        return er == hipSuccess ? VX_OK : fail(VX_ERR_DEVICE, "upload failed: %s", hipGetErrorString(er));
    };
    int32_t rc = VX_OK;
#define UP(x)                          \
    do {                               \
        rc = (x);                      \
        if (rc != VX_OK) return fail_free(rc); \
    } while (0)
    UP(up_d(G.pos, desc->positions, 3));
    UP(up_d(G.pred, desc->positions, 3));
    UP(up_d(G.prev, desc->positions, 3));
    UP(up_d(G.vel, desc->velocities, 3));
    UP(up_d(G.drift, nullptr, 3));
    UP(up_d(G.rot, desc->rotations, 4));
    {
        std::vector<float> tg(3 * N);
        for (size_t k = 0; k < 3 * N; ++k) tg[k] = desc->targets ? desc->targets[k] : 0.0f;
        for (size_t k = 0; k < 3 * N; ++k)
            if (std::isnan(tg[k])) tg[k] = 0.0f;
        UP(up_d(G.targets, tg.data(), 3));
    }
    UP(up_f(G.mass, desc->mass, 1, 0));
    UP(up_f(G.young, desc->young, 1, 0));
    UP(up_f(G.young_perp, desc->young_perp, 1, 0));
    UP(up_f(G.shear, desc->shear, 1, 0));
    UP(up_f(G.solid, desc->solid, 1, 1));
    UP(up_f(G.fiber, desc->fibers, 3, 0));
    UP(up_f(G.aniso, desc->anisotropy, 1, 0));
    UP(up_f(G.dtemp, desc->delta_temp, 1, 0));
    UP(up_f(G.a_par, desc->alpha_par, 1, 0));
    UP(up_f(G.a_perp, desc->alpha_perp, 1, 0));
    UP(up_f(G.loads, desc->loads, 3, 0));
    UP(up_bond(G.damage, desc->damage, 0.0f));
    UP(up_bond(G.kscale, desc->kscale, 1.0f));
    UP(up_bond_vec(G.offset, desc->formation_offset));
    UP(up_bond_vec(G.plastic, desc->plastic_strain));
#undef UP
    if (hipMemcpy(G.kinematic, desc->kinematic, N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(G.bonds, bonds.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(G.tmask, tmask.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(G.mask, mask.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(G.xa, 0, 24 * N) != hipSuccess || hipMemset(G.xb, 0, 24 * N) != hipSuccess ||
        hipMemset(G.d, 0, 12 * N) != hipSuccess || hipMemset(G.dd, 0, 24 * N) != hipSuccess ||
        hipMemset(G.st, 0, sizeof(GridState)) != hipSuccess)
        return fail_free(fail(VX_ERR_DEVICE, "scene upload failed"));
    if (hipHostMalloc((void **)&sc->pinned, 3 * sizeof(GridState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&sc->host_partials, 2 * sizeof(double) * sc->grid, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&sc->ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sc->ev[1], hipEventDisableTiming) != hipSuccess)
        return fail_free(fail(VX_ERR_DEVICE, "scene host resources failed"));
    (void)ctx;
    *out = sc;
    return VX_OK;
}

extern "C" int32_t vx_scene_destroy(vx_scene *sc) {
    clear_error();
    if (!sc) return VX_OK;
    Lock lock;
    (void)hipStreamSynchronize(context().stream);
    if (sc->inner32) (void)hipGraphExecDestroy(sc->inner32);
    if (sc->inner64) (void)hipGraphExecDestroy(sc->inner64);
    for (auto &e : sc->ev)
        if (e) (void)hipEventDestroy(e);
    if (sc->pinned) (void)hipHostFree(sc->pinned);
    if (sc->host_partials) (void)hipHostFree(sc->host_partials);
    if (sc->alloc) (void)hipFree(sc->alloc);
    delete sc;
    return VX_OK;
}

extern "C" int32_t vx_scene_predict(vx_scene *sc, double h, const double *g) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    Lock lock;
    const double gx = g ? g[0] : 0.0, gy = g ? g[1] : 0.0, gz = g ? g[2] : 0.0;
    hipLaunchKernelGGL(grid_predict, dim3(blocks_for(sc->s.g.n)), dim3(kBlock), 0, context().stream, sc->s, h, gx, gy, gz);
    VX_HIP(hipGetLastError());
    return VX_OK;
}

extern "C" int32_t vx_scene_fit_rotations(vx_scene *sc, double *out_max_angle) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    Lock lock;
    Context &ctx = context();
    VX_HIP(hipMemsetAsync(&sc->s.st->angle_bits, 0, sizeof(uint64_t), ctx.stream));
    hipLaunchKernelGGL(grid_fit, dim3(blocks_for(sc->s.g.n)), dim3(kBlock), 0, ctx.stream, sc->s);
    VX_HIP(hipGetLastError());
    if (out_max_angle) {
        GridState st;
        if (int32_t c = read_state(sc, &st)) return c;
        double a;
        std::memcpy(&a, &st.angle_bits, sizeof a);
        *out_max_angle = a;
    }
    return VX_OK;
}

extern "C" int32_t vx_scene_assemble(vx_scene *sc, double h) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    if (!(h > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "substep length must be positive");
    Lock lock;
    Context &ctx = context();
    hipLaunchKernelGGL(grid_bonds, dim3(blocks_for(sc->s.g.n)), dim3(kBlock), 0, ctx.stream, sc->s);
    VX_HIP(hipGetLastError());
    hipLaunchKernelGGL(grid_rows, dim3(blocks_for(sc->s.g.n)), dim3(kBlock), 0, ctx.stream, sc->s, h);
    VX_HIP(hipGetLastError());
    return VX_OK;
}

extern "C" int32_t vx_scene_solve(vx_scene *sc, const vx_solve_params *prm, vx_solve_report *rep) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    if (!prm || prm->max_inner == 0 || !(prm->tolerance > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "invalid solve parameters");
    Lock lock;
    if (rep) std::memset(rep, 0, sizeof *rep);
    if (prm->mode == VX_SOLVE_F64) return solve_t<double>(sc, prm, rep);
    return solve_t<float>(sc, prm, rep);
}

extern "C" int32_t vx_scene_apply(vx_scene *sc) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    Lock lock;
    hipLaunchKernelGGL(grid_apply, dim3(blocks_for(sc->s.g.n)), dim3(kBlock), 0, context().stream, sc->s,
                       (const double *)(sc->x_in_a ? sc->s.xa : sc->s.xb));
    VX_HIP(hipGetLastError());
    return VX_OK;
}

extern "C" int32_t vx_scene_finish_substep(vx_scene *sc, double h) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    Lock lock;
    hipLaunchKernelGGL(grid_finish, dim3(blocks_for(sc->s.g.n)), dim3(kBlock), 0, context().stream, sc->s, h);
    VX_HIP(hipGetLastError());
    return VX_OK;
}

static int32_t download_planes(const double *dev, size_t n, int comps, double *out) {
    std::vector<double> buf(comps * n);
    VX_HIP(hipMemcpyAsync(buf.data(), dev, sizeof(double) * comps * n, hipMemcpyDeviceToHost, context().stream));
    VX_HIP(hipStreamSynchronize(context().stream));
    for (size_t v = 0; v < n; ++v)
        for (int q = 0; q < comps; ++q) out[comps * v + q] = buf[q * n + v];
    return VX_OK;
}

extern "C" int32_t vx_scene_download(vx_scene *sc, double *positions, double *rotations, double *velocities, double *correction) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    Lock lock;
    const size_t n = sc->s.g.n;
    if (positions)
        if (int32_t c = download_planes(sc->s.pos, n, 3, positions)) return c;
    if (rotations)
        if (int32_t c = download_planes(sc->s.rot, n, 4, rotations)) return c;
    if (velocities)
        if (int32_t c = download_planes(sc->s.vel, n, 3, velocities)) return c;
    if (correction)
        if (int32_t c = download_planes(sc->x_in_a ? sc->s.xa : sc->s.xb, n, 3, correction)) return c;
    return VX_OK;
}

// Reference row packing (src/solver.rs:1652-1672, 1731-1777) of the assembled
// structured operator: slots in incidence order (-x, +x, -y, +y, -z, +z with
// absent bonds compacted out), off-diagonal blocks -P_i k P_j, zero blocks and
// EMPTY for inactive neighbours.
extern "C" int32_t vx_scene_export_rows(vx_scene *sc, uint32_t *active_out, uint32_t *neighbors, double *blocks, double *diagonal,
                                        double *rhs_out, uint32_t *out_active_count) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    Lock lock;
    const DevScene &G = sc->s;
    const size_t n = G.g.n;
    std::vector<double> k(18 * n), d(6 * n), rhs(3 * n);
    std::vector<uint8_t> mask(n), bonds(n), kin(n);
    Context &ctx = context();
    VX_HIP(hipMemcpyAsync(k.data(), G.k64, 8 * 18 * n, hipMemcpyDeviceToHost, ctx.stream));
    VX_HIP(hipMemcpyAsync(d.data(), G.d64, 8 * 6 * n, hipMemcpyDeviceToHost, ctx.stream));
    VX_HIP(hipMemcpyAsync(rhs.data(), G.rhs, 8 * 3 * n, hipMemcpyDeviceToHost, ctx.stream));
    VX_HIP(hipMemcpyAsync(mask.data(), G.mask, n, hipMemcpyDeviceToHost, ctx.stream));
    VX_HIP(hipMemcpyAsync(bonds.data(), G.bonds, n, hipMemcpyDeviceToHost, ctx.stream));
    VX_HIP(hipMemcpyAsync(kin.data(), G.kinematic, n, hipMemcpyDeviceToHost, ctx.stream));
    VX_HIP(hipStreamSynchronize(ctx.stream));
    std::vector<uint32_t> aidx(n, VX_EMPTY);
    uint32_t na = 0;
    for (size_t v = 0; v < n; ++v)
        if (mask[v]) aidx[v] = na++;
    const uint32_t nx = G.g.nx, ny = G.g.ny;
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
        uint32_t e = 0;
        if (neighbors)
            for (int s = 0; s < 6; ++s) neighbors[6 * size_t(slot) + s] = VX_EMPTY;
        if (blocks)
            for (int s = 0; s < 54; ++s) blocks[54 * size_t(slot) + s] = 0.0;
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
                // Constraint enters the system (one endpoint has a free axis).
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
    (void)kin;
    if (out_active_count) *out_active_count = na;
    return VX_OK;
}

extern "C" int32_t vx_scene_time_spmv(vx_scene *sc, uint32_t reps, double *out_us) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    if (reps == 0 || !out_us) return fail(VX_ERR_INVALID_ARGUMENT, "reps must be positive");
    Lock lock;
    Context &ctx = context();
    DevScene &G = sc->s;
    // The timed launches run as first inner iterations (k = 0, nonzero
    // residual partials, not done); they leave the solve vectors' contents
    // unspecified.
    GridState init;
    std::memset(&init, 0, sizeof init);
    init.inner_tol = 0.0;
    init.max_iter = 1u << 30;
    sc->pinned[2] = init;
    VX_HIP(hipMemcpyAsync(G.st, &sc->pinned[2], sizeof init, hipMemcpyHostToDevice, ctx.stream));
    std::vector<double> ones(2 * size_t(sc->grid), 1.0);
    VX_HIP(hipMemcpyAsync(G.partials + size_t(kRz) * sc->grid, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice,
                          ctx.stream));
    std::vector<hipEvent_t> ev(2 * reps);
    for (auto &e : ev) VX_HIP(hipEventCreate(&e));
    for (uint32_t k = 0; k < reps; ++k) {
        VX_HIP(hipEventRecord(ev[2 * k], ctx.stream));
        hipLaunchKernelGGL(grid_spmv<float>, dim3(sc->grid), dim3(kBlock), 0, ctx.stream, G, sc->per_xcd, sc->ntiles, G.z,
                           (k % 2) ? G.p1 : G.p0, (k % 2) ? G.p0 : G.p1, G.ap);
        VX_HIP(hipEventRecord(ev[2 * k + 1], ctx.stream));
    }
    VX_HIP(hipStreamSynchronize(ctx.stream));
    double total = 0.0;
    for (uint32_t k = 0; k < reps; ++k) {
        float ms = 0.0f;
        VX_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        total += ms;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    *out_us = 1000.0 * total / reps;
    return VX_OK;
}

extern "C" int32_t vx_scene_step(vx_scene *sc, const vx_step_params *prm, vx_step_report *rep) {
    clear_error();
    if (int32_t c = check_scene(sc)) return c;
    if (!prm || prm->substeps == 0 || prm->corotation_iterations == 0 || !(prm->fixed_dt > 0.0))
        return fail(VX_ERR_INVALID_ARGUMENT, "invalid step parameters");
    Lock lock;
    vx_step_report out;
    std::memset(&out, 0, sizeof out);
    out.active_rows = sc->active;
    const auto t0 = std::chrono::steady_clock::now();
    const double h = prm->fixed_dt / prm->substeps;
    int32_t status = VX_OK;
    for (uint32_t sub = 0; sub < prm->substeps && status == VX_OK; ++sub) {
        if ((status = vx_scene_predict(sc, h, prm->gravity))) break;
        if ((status = vx_scene_fit_rotations(sc, nullptr))) break;
        for (uint32_t pass = 0; pass < prm->corotation_iterations; ++pass) {
            if ((status = vx_scene_assemble(sc, h))) break;
            vx_solve_report sr;
            status = vx_scene_solve(sc, &prm->solve, &sr);
            out.solves += 1;
            out.inner_iterations += sr.inner_iterations;
            out.outer_sweeps += sr.outer_sweeps;
            out.last_relative_residual = sr.relative_residual;
            if (status) break;
            if ((status = vx_scene_apply(sc))) break;
            if (pass + 1 == prm->corotation_iterations) break;
            double angle = 0.0;
            if ((status = vx_scene_fit_rotations(sc, &angle))) break;
            out.last_angle = angle;
            if (angle <= prm->corotation_tolerance && sr.relative_residual <= prm->solve.tolerance) break;
        }
        if (status) break;
        status = vx_scene_finish_substep(sc, h);
    }
    if (status == VX_OK) {
        hipError_t e = hipStreamSynchronize(context().stream);
        if (e != hipSuccess) status = fail(VX_ERR_DEVICE, "step synchronize failed: %s", hipGetErrorString(e));
    }
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep) *rep = out;
    return status;
}
