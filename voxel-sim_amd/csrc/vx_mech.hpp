// FP64 device mechanics of the corotational bond model: small vector,
// matrix and quaternion algebra, bond frames, anisotropic compliance and the
// rotated bond stiffness. Restates src/solver.rs:27-110 and :1325-1345.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vx {

constexpr double kCell = 0.01;  // src/scene.rs:10 (REV edge, metres)

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
// Scale by the reciprocal length when it is finite and positive, else zero.
__device__ __forceinline__ d3 normalize_or_zero(d3 a) {
    double rcp = 1.0 / sqrt(dot3(a, a));
    if (isfinite(rcp) && rcp > 0.0) return a * rcp;
    return {0.0, 0.0, 0.0};
}

struct m3 {  // columns
    d3 c0, c1, c2;
};
__device__ __forceinline__ d3 col(const m3 &m, int i) { return i == 0 ? m.c0 : (i == 1 ? m.c1 : m.c2); }
__device__ __forceinline__ d3 mul(const m3 &m, d3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }
__device__ __forceinline__ d3 mul_t(const m3 &m, d3 v) { return {dot3(m.c0, v), dot3(m.c1, v), dot3(m.c2, v)}; }
__device__ __forceinline__ double det3(const m3 &m) { return dot3(m.c2, cross3(m.c0, m.c1)); }
__device__ __forceinline__ m3 transpose3(const m3 &m) {
    return {{m.c0.x, m.c1.x, m.c2.x}, {m.c0.y, m.c1.y, m.c2.y}, {m.c0.z, m.c1.z, m.c2.z}};
}
// Inverse by cross products of column pairs (the rows of the inverse).
__device__ __forceinline__ m3 inverse3m(const m3 &m) {
    d3 t0 = cross3(m.c1, m.c2), t1 = cross3(m.c2, m.c0), t2 = cross3(m.c0, m.c1);
    double inv = 1.0 / dot3(m.c2, t2);
    return transpose3(m3{t0 * inv, t1 * inv, t2 * inv});
}

struct q4 {  // x, y, z, w
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
// Mike Day's method ("Converting a Rotation Matrix to a Quaternion", 2015),
// the branch structure glam's Quat::from_mat3 uses: the sign of m22 and of
// m00 -/+ m11 picks the largest of w^2, x^2, y^2, z^2 for the square root.
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
// Half-way slerp with the shortest-arc flip; normalized lerp above 0.9995
// (the reference's Quat::slerp(.., 0.5), src/solver.rs:1326).
__device__ __forceinline__ q4 slerp_half(q4 a, q4 b) {
    double d = qdot(a, b);
    if (d < 0.0) {
        b = {-b.x, -b.y, -b.z, -b.w};
        d = -d;
    }
    if (d > 0.9995)
        return qnormalize({a.x + (b.x - a.x) * 0.5, a.y + (b.y - a.y) * 0.5, a.z + (b.z - a.z) * 0.5, a.w + (b.w - a.w) * 0.5});
    double th = acos(d);
    double s1 = sin(th * 0.5), inv = 1.0 / sin(th);
    return {(a.x + b.x) * s1 * inv, (a.y + b.y) * s1 * inv, (a.z + b.z) * s1 * inv, (a.w + b.w) * s1 * inv};
}
// basis(axis, q): rotated e_axis, e_axis+1, e_axis+2 (src/solver.rs:27-34).
__device__ __forceinline__ m3 basis(int axis, q4 q) {
    const d3 e[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    return {qrotate(q, e[axis]), qrotate(q, e[(axis + 1) % 3]), qrotate(q, e[(axis + 2) % 3])};
}

// Symmetric 3x3 as (xx, xy, xz, yy, yz, zz).
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

// Material constants of one REV as the solve reads them.
struct Rev {
    double young, young_perp, shear, solid, aniso;
    d3 fiber;  // body frame
    double dtemp, a_par, a_perp;
};

// tensor_compliance (src/solver.rs:37-63) in the bond frame F for a voxel
// rotated by rv; false when the compliance is invalid (the reference's ZERO).
__device__ __forceinline__ bool compliance(const Rev &m, const m3 &F, q4 rv, double (&C)[6]) {
    double solid = fmin(fmax(m.solid, 1e-4), 1.0);
    d3 fib = mul_t(F, qrotate(rv, m.fiber));
    double tr = m.young + m.aniso * (m.young_perp - m.young);
    double nrm = tr + (m.young - tr) * fib.x * fib.x;
    d3 dg = mk(nrm, m.shear, m.shear) * solid;
    if (!(isfinite(dg.x) && isfinite(dg.y) && isfinite(dg.z)) || fmin(fmin(dg.x, dg.y), dg.z) <= 0.0) return false;
    d3 inv{1.0 / dg.x, 1.0 / dg.y, 1.0 / dg.z};
    d3 sf{fib.x * inv.x, fib.y * inv.y, fib.z * inv.z};
    double alpha = m.shear * m.aniso * 0.2;
    double c = alpha / (1.0 + alpha * dot3(fib, sf));
    C[0] = inv.x - sf.x * sf.x * c;
    C[1] = -sf.x * sf.y * c;
    C[2] = -sf.x * sf.z * c;
    C[3] = inv.y - sf.y * sf.y * c;
    C[4] = -sf.y * sf.z * c;
    C[5] = inv.z - sf.z * sf.z * c;
    return true;
}

// Free thermal strain of one endpoint along the bond direction
// (src/solver.rs:80-92).
__device__ __forceinline__ d3 endpoint_expansion(const Rev &m, q4 rv, d3 dir) {
    d3 fib = normalize_or_zero(qrotate(rv, m.fiber));
    d3 ad = dir * m.a_perp + fib * ((m.a_par - m.a_perp) * dot3(fib, dir));
    return ad * m.dtemp;
}

// Bond stiffness in the bond frame F, K = (Ca + Cb)^-1 * 2*AREA/CELL * scale
// (src/solver.rs:65-73), symmetric by cofactors; false (K = 0) when either
// compliance is invalid.
__device__ __forceinline__ bool local_stiffness(const m3 &F, const Rev &mi, q4 qi, const Rev &mj, q4 qj, double scale,
                                                double (&K)[6]) {
    double Ca[6], Cb[6];
    for (int q = 0; q < 6; ++q) K[q] = 0.0;
    if (!(compliance(mi, F, qi, Ca) && compliance(mj, F, qj, Cb))) return false;
    double s0 = Ca[0] + Cb[0], s1 = Ca[1] + Cb[1], s2 = Ca[2] + Cb[2];
    double s3 = Ca[3] + Cb[3], s4 = Ca[4] + Cb[4], s5 = Ca[5] + Cb[5];
    double c00 = s3 * s5 - s4 * s4, c01 = s2 * s4 - s1 * s5, c02 = s1 * s4 - s2 * s3;
    double c11 = s0 * s5 - s2 * s2, c12 = s1 * s2 - s0 * s4, c22 = s0 * s3 - s1 * s1;
    double det = s0 * c00 + s1 * c01 + s2 * c02;
    double sc = (2.0 * kCell * kCell / kCell) / det * scale;
    K[0] = c00 * sc;
    K[1] = c01 * sc;
    K[2] = c02 * sc;
    K[3] = c11 * sc;
    K[4] = c12 * sc;
    K[5] = c22 * sc;
    return true;
}

// Thermal rest length vector of the bond in its own frame F
// (bond_thermal_rest, src/solver.rs:78-95), endpoints rotated by qi, qj.
__device__ __forceinline__ d3 thermal_rest_local(const m3 &F, const Rev &mi, q4 qi, const Rev &mj, q4 qj) {
    d3 dir = normalize_or_zero(F.c0);
    return mul_t(F, dir * kCell + (endpoint_expansion(mi, qi, dir) + endpoint_expansion(mj, qj, dir)) * (0.5 * kCell));
}

// Rotated bond stiffness k = F (K * scale) F^T (symmetric, 6 values) and the
// rest vector rest = F (thermal + offset + plastic CELL) of the +axis bond
// i -> j (src/solver.rs:65-95, 1325-1345). scale = kscale * max(1 - damage,
// 0.001). k is zero when either compliance is invalid.
__device__ __forceinline__ void bond_constraint(int axis, const Rev &mi, q4 qi, const Rev &mj, q4 qj, double scale,
                                                d3 offset, d3 plastic, double (&k)[6], d3 &rest) {
    const m3 F = basis(axis, slerp_half(qi, qj));
    for (int q = 0; q < 6; ++q) k[q] = 0.0;
    double Ks[6];
    if (local_stiffness(F, mi, qi, mj, qj, scale, Ks)) {
        double K[3][3] = {{Ks[0], Ks[1], Ks[2]}, {Ks[1], Ks[3], Ks[4]}, {Ks[2], Ks[4], Ks[5]}};
        double Fm[3][3] = {{F.c0.x, F.c1.x, F.c2.x}, {F.c0.y, F.c1.y, F.c2.y}, {F.c0.z, F.c1.z, F.c2.z}};
        double T[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) T[r][c] = Fm[r][0] * K[0][c] + Fm[r][1] * K[1][c] + Fm[r][2] * K[2][c];
        int q = 0;
        for (int r = 0; r < 3; ++r)
            for (int c = r; c < 3; ++c) k[q++] = T[r][0] * Fm[c][0] + T[r][1] * Fm[c][1] + T[r][2] * Fm[c][2];
    }
    rest = mul(F, thermal_rest_local(F, mi, qi, mj, qj) + offset + plastic * kCell);
}

// ---- fracture law (src/fracture.rs) -----------------------------------------
// Per-REV strength row (Material fields, src/material.rs:20-63).
enum : int {
    S_TENSION, S_TENSION_PERP, S_COMPRESSION, S_COMPRESSION_PERP, S_SHEAR, S_SHEAR_PERP, S_YIELD, S_HARDENING,
    S_PLASTIC_FAILURE, S_MODE, S_GC, S_GC_PERP, kStrength
};
enum : int { M_BRITTLE = 0, M_DUCTILE = 1, M_MIXED = 2 };
// Per-bond history beyond damage and plastic strain (Bond, src/bond.rs:4-16).
enum : int { H_EQ_PLASTIC, H_PLASTIC_WORK, H_FATIGUE, H_PREVIOUS, H_REVERSAL, H_INCREMENT, kHistory };

struct Limits {
    double tension, compression, shear, yield, hardening, plastic_failure, ductility, gc;
};
struct FractureCfg {
    double damage_rate, energy_release, fatigue_cycles, fatigue_exponent, fatigue_threshold, heat_fraction;
};
// plastic_heat, stored_plastic, surface_work, damage_heat, released
struct Dissipation {
    double e[5];
};

// directional (src/fracture.rs:73-75).
__device__ __forceinline__ double directional(double par, double perp, double alignment, double anisotropy) {
    const double a = fmin(fmax(fabs(alignment), 0.0), 1.0);
    return par + anisotropy * (1.0 - a) * (perp - par);
}

// One endpoint's limits (src/fracture.rs:76-127); s is its strength row.
__device__ __forceinline__ Limits endpoint_limits(const float (&s)[kStrength], double alignment, double anisotropy) {
    Limits l;
    const int mode = int(s[S_MODE]);
    l.ductility = mode == M_BRITTLE ? 0.0
                  : mode == M_DUCTILE ? 1.0
                                      : (1.0 - anisotropy) * 0.5 + anisotropy * (alignment * alignment);
    l.tension = directional(s[S_TENSION], s[S_TENSION_PERP], alignment, anisotropy);
    l.compression = directional(s[S_COMPRESSION], s[S_COMPRESSION_PERP], alignment, anisotropy);
    l.shear = directional(s[S_SHEAR], s[S_SHEAR_PERP], alignment, anisotropy);
    l.yield = s[S_YIELD];
    l.hardening = s[S_HARDENING];
    l.plastic_failure = s[S_PLASTIC_FAILURE];
    l.gc = directional(s[S_GC], s[S_GC_PERP], alignment, anisotropy);
    return l;
}

// pair_limits (src/fracture.rs:128-133): the weaker endpoint per field,
// scaled by the smaller solid fraction squared.
__device__ __forceinline__ Limits pair_limits(const Limits &a, const Limits &b, double scale) {
    return Limits{fmin(a.tension, b.tension) * scale,
                  fmin(a.compression, b.compression) * scale,
                  fmin(a.shear, b.shear) * scale,
                  fmin(a.yield, b.yield) * scale,
                  fmin(a.hardening, b.hardening) * scale,
                  fmin(a.plastic_failure, b.plastic_failure),
                  fmin(a.ductility, b.ductility),
                  fmin(a.gc, b.gc) * scale};
}

__device__ __forceinline__ d3 div3(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }

// Symmetric 3x3 (xx, xy, xz, yy, yz, zz) times a vector.
__device__ __forceinline__ d3 sym_mul(const double (&k)[6], d3 u) {
    return {k[0] * u.x + k[1] * u.y + k[2] * u.z, k[1] * u.x + k[3] * u.y + k[4] * u.z, k[2] * u.x + k[4] * u.y + k[5] * u.z};
}

// Elastic energy of a bond (src/fracture.rs:186-188).
__device__ __forceinline__ double bond_energy(d3 u, const double (&k)[6], double damage) {
    return fmax(0.5 * (1.0 - damage) * dot3(u, sym_mul(k, u)), 0.0);
}

// Fatigue counting (src/fracture.rs:190-207): only reversals of the signed
// stress ratio consume life.
__device__ __forceinline__ void fatigue(double (&h)[kHistory], double signed_ratio, const FractureCfg &c) {
    const double delta = signed_ratio - h[H_PREVIOUS];
    if (fabs(delta) > 1e-6) {
        if (h[H_INCREMENT] * delta < 0.0) {
            const double amplitude = fabs(h[H_PREVIOUS] - h[H_REVERSAL]) * 0.5;
            if (amplitude > c.fatigue_threshold) {
                const double level = fmin((amplitude - c.fatigue_threshold) / (1.0 - c.fatigue_threshold), 1.0);
                h[H_FATIGUE] = fmin(h[H_FATIGUE] + 0.5 * pow(level, c.fatigue_exponent) / c.fatigue_cycles, 1.0);
            }
            h[H_REVERSAL] = h[H_PREVIOUS];
        }
        h[H_INCREMENT] = delta;
    }
    h[H_PREVIOUS] = signed_ratio;
}

// Constitutive update of one bond (src/fracture.rs:210-292): plastic return
// mapping along the trial traction, the stress ratio against the limits,
// fatigue, damage growth, and the split of the energy lost. u_total is the
// bond extension in its frame, k the material stiffness (undamaged). Returns
// true when the bond breaks (damage reaches one).
__device__ __forceinline__ bool fracture_update(d3 &plastic, double &damage, double (&h)[kHistory], d3 u_total,
                                                const double (&k)[6], const Limits &l, double dt, const FractureCfg &c,
                                                Dissipation &r) {
    for (int q = 0; q < 5; ++q) r.e[q] = 0.0;
    const double area = kCell * kCell;
    const double old_damage = damage;
    d3 u = u_total - plastic * kCell;
    const double before = bond_energy(u, k, old_damage);
    const d3 trial = div3(sym_mul(k, u), area);
    if (l.ductility > 1e-5) {
        const double equivalent = sqrt(dot3(trial, trial));
        const double yield_now = l.yield + l.hardening * h[H_EQ_PLASTIC];
        if (equivalent > yield_now) {
            const d3 direction = div3(trial, equivalent);
            double kinv[6];
            sym_inverse(k, kinv);
            const d3 cdir = div3(sym_mul(kinv, direction * area), kCell);
            const double modulus = 1.0 / fmax(sqrt(dot3(cdir, cdir)), 1e-20);
            const double dep = (equivalent - yield_now) / (modulus + l.hardening) * l.ductility;
            plastic = plastic + normalize_or_zero(cdir) * dep;
            h[H_EQ_PLASTIC] += dep;
            u = u_total - plastic * kCell;
            const double lost = fmax(before - bond_energy(u, k, old_damage), 0.0);
            h[H_PLASTIC_WORK] += lost;
            r.e[0] = lost * c.heat_fraction;
            r.e[1] = lost - r.e[0];
        }
    }
    const d3 stress = div3(sym_mul(k, u), area);
    const double normal_limit = stress.x >= 0.0 ? l.tension : l.compression;
    const double sx = stress.x / fmax(normal_limit, 1.0);
    const double sh = fmax(l.shear, 1.0);
    const double ratio = sqrt(sx * sx + (stress.y * stress.y + stress.z * stress.z) / (sh * sh));
    double principal;
    if (fabs(stress.x) >= fmax(fabs(stress.y), fabs(stress.z)))
        principal = stress.x;
    else if (fabs(stress.y) >= fabs(stress.z))
        principal = stress.y;
    else
        principal = stress.z;
    fatigue(h, copysign(ratio, principal), c);
    if (ratio >= 1.0) {
        if (l.ductility <= 1e-5)
            damage = 1.0;
        else
            damage = fmin(damage + c.damage_rate * dt * fmin(ratio, 10.0), 1.0);
    }
    if (l.ductility > 1e-5 && h[H_EQ_PLASTIC] >= fmax(l.plastic_failure, 1e-6)) damage = fmin(damage + c.damage_rate * dt, 1.0);
    if (h[H_FATIGUE] >= 1.0) damage = 1.0;
    const double lost = fmax(bond_energy(u, k, old_damage) - bond_energy(u, k, damage), 0.0);
    const double budget = l.gc * area * (damage - old_damage);
    r.e[2] = fmin(lost, budget);
    const double available = lost - r.e[2];
    r.e[4] = available * c.energy_release;
    r.e[3] = available - r.e[4];
    return damage >= 1.0;
}

}  // namespace vx
