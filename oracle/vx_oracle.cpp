// CPU oracle for the corotational elastic PCG solve. TEST INFRASTRUCTURE ONLY:
// what the tests compare the HIP path with, and the timed CPU baseline; never
// part of the product path.
// Each function cites the reference file:line it restates (paths relative to
// the reference repository). See vx_oracle.h for the API contract.
#include "vx_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kCell = 0.01;  // src/scene.rs:10
constexpr uint32_t kEmpty = VXO_EMPTY;

// ---------------------------------------------------------------------------
// Small linear algebra in the reference's (glam) conventions: Mat3 is stored
// by columns, matrix*vector accumulates columns in x, y, z order.
// ---------------------------------------------------------------------------
template <class T>
struct V3 {
    T x, y, z;
    T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> T max_abs_element(V3<T> a) {
    return std::max(std::max(std::fabs(a.x), std::fabs(a.y)), std::fabs(a.z));
}
template <class T> bool finite3(V3<T> a) {
    return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z);
}
// glam normalize_or_zero: scale by the reciprocal length when it is finite
// and positive, else the zero vector.
template <class T> V3<T> normalize_or_zero(V3<T> a) {
    T rcp = T(1) / std::sqrt(dot(a, a));
    if (std::isfinite(rcp) && rcp > T(0)) return a * rcp;
    return {T(0), T(0), T(0)};
}

template <class T>
struct M3 {
    V3<T> c[3];  // columns
    static M3 zero() { return M3{{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}}; }
    static M3 identity() { return M3{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}}; }
    static M3 diag(V3<T> d) { return M3{{{d.x, 0, 0}, {0, d.y, 0}, {0, 0, d.z}}}; }
    T at(int r, int col) const { return c[col][r]; }
};
template <class T> V3<T> operator*(const M3<T> &m, V3<T> v) {
    V3<T> r = m.c[0] * v.x;
    r = r + m.c[1] * v.y;
    r = r + m.c[2] * v.z;
    return r;
}
template <class T> M3<T> operator*(const M3<T> &a, const M3<T> &b) {
    return M3<T>{{a * b.c[0], a * b.c[1], a * b.c[2]}};
}
template <class T> M3<T> operator+(const M3<T> &a, const M3<T> &b) {
    return M3<T>{{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}};
}
template <class T> M3<T> operator-(const M3<T> &a, const M3<T> &b) {
    return M3<T>{{a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]}};
}
template <class T> M3<T> operator-(const M3<T> &a) { return M3<T>{{-a.c[0], -a.c[1], -a.c[2]}}; }
template <class T> M3<T> operator*(const M3<T> &a, T s) {
    return M3<T>{{a.c[0] * s, a.c[1] * s, a.c[2] * s}};
}
template <class T> M3<T> transpose(const M3<T> &m) {
    return M3<T>{{{m.c[0].x, m.c[1].x, m.c[2].x},
                  {m.c[0].y, m.c[1].y, m.c[2].y},
                  {m.c[0].z, m.c[1].z, m.c[2].z}}};
}
template <class T> T determinant(const M3<T> &m) { return dot(m.c[2], cross(m.c[0], m.c[1])); }
// Adjugate (cross-product) inverse: rows of the inverse are the cross
// products of column pairs divided by the determinant.
template <class T> M3<T> inverse(const M3<T> &m) {
    V3<T> t0 = cross(m.c[1], m.c[2]);
    V3<T> t1 = cross(m.c[2], m.c[0]);
    V3<T> t2 = cross(m.c[0], m.c[1]);
    T det = dot(m.c[2], t2);
    T inv = T(1) / det;
    return transpose(M3<T>{{t0 * inv, t1 * inv, t2 * inv}});
}
template <class T> bool mat_is_zero(const M3<T> &m) {
    for (int i = 0; i < 3; ++i)
        for (int r = 0; r < 3; ++r)
            if (m.c[i][r] != T(0)) return false;
    return true;
}

template <class T>
struct Q4 {
    T x, y, z, w;
};
template <class T> T qdot(Q4<T> a, Q4<T> b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
template <class T> Q4<T> qnormalize(Q4<T> q) {
    T len = std::sqrt(qdot(q, q));
    T rcp = T(1) / len;
    return {q.x * rcp, q.y * rcp, q.z * rcp, q.w * rcp};
}
// Rotate a vector by a unit quaternion: v(w^2 - b.b) + b 2(v.b) + (b x v) 2w.
template <class T> V3<T> qrotate(Q4<T> q, V3<T> v) {
    V3<T> b{q.x, q.y, q.z};
    T b2 = dot(b, b);
    return v * (q.w * q.w - b2) + b * (dot(v, b) * T(2)) + cross(b, v) * (q.w * T(2));
}
// Rotation matrix of a unit quaternion (standard closed form).
template <class T> M3<T> mat_from_quat(Q4<T> q) {
    T x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    T xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    T yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    T wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    return M3<T>{{{T(1) - (yy + zz), xy + wz, xz - wy},
                  {xy - wz, T(1) - (xx + zz), yz + wx},
                  {xz + wy, yz - wx, T(1) - (xx + yy)}}};
}
// Quaternion of a rotation matrix: Mike Day's method ("Converting a Rotation
// Matrix to a Quaternion", 2015), glam Quat::from_mat3's branch structure:
// the signs of m22 and m00 -/+ m11 pick the largest of w^2, x^2, y^2, z^2
// for a stable square root.
template <class T> Q4<T> quat_from_mat(const M3<T> &m) {
    T m00 = m.c[0].x, m01 = m.c[0].y, m02 = m.c[0].z;
    T m10 = m.c[1].x, m11 = m.c[1].y, m12 = m.c[1].z;
    T m20 = m.c[2].x, m21 = m.c[2].y, m22 = m.c[2].z;
    if (m22 <= T(0)) {
        T dif10 = m11 - m00;
        T omm22 = T(1) - m22;
        if (dif10 <= T(0)) {
            T four_xsq = omm22 - dif10;
            T inv4x = T(0.5) / std::sqrt(four_xsq);
            return {four_xsq * inv4x, (m01 + m10) * inv4x, (m02 + m20) * inv4x, (m12 - m21) * inv4x};
        }
        T four_ysq = omm22 + dif10;
        T inv4y = T(0.5) / std::sqrt(four_ysq);
        return {(m01 + m10) * inv4y, four_ysq * inv4y, (m12 + m21) * inv4y, (m20 - m02) * inv4y};
    }
    T sum10 = m11 + m00;
    T opm22 = T(1) + m22;
    if (sum10 <= T(0)) {
        T four_zsq = opm22 - sum10;
        T inv4z = T(0.5) / std::sqrt(four_zsq);
        return {(m02 + m20) * inv4z, (m12 + m21) * inv4z, four_zsq * inv4z, (m01 - m10) * inv4z};
    }
    T four_wsq = opm22 + sum10;
    T inv4w = T(0.5) / std::sqrt(four_wsq);
    return {(m12 - m21) * inv4w, (m20 - m02) * inv4w, (m01 - m10) * inv4w, four_wsq * inv4w};
}
// Half-way spherical interpolation with the shortest-arc sign flip; near-
// parallel pairs use normalized lerp (threshold 0.9995), as the reference's
// Quat::slerp does (src/solver.rs:1326, :101).
template <class T> Q4<T> slerp_half(Q4<T> a, Q4<T> b) {
    T d = qdot(a, b);
    if (d < T(0)) {
        b = {-b.x, -b.y, -b.z, -b.w};
        d = -d;
    }
    const T s = T(0.5);
    if (d > T(0.9995)) {
        Q4<T> q{a.x + (b.x - a.x) * s, a.y + (b.y - a.y) * s, a.z + (b.z - a.z) * s,
                a.w + (b.w - a.w) * s};
        return qnormalize(q);
    }
    T theta = std::acos(d);
    T s1 = std::sin(theta * (T(1) - s));
    T s2 = std::sin(theta * s);
    T inv = T(1) / std::sin(theta);
    return {(a.x * s1 + b.x * s2) * inv, (a.y * s1 + b.y * s2) * inv, (a.z * s1 + b.z * s2) * inv,
            (a.w * s1 + b.w * s2) * inv};
}

// basis (src/solver.rs:27-34): rotated axis, next axis, next-next axis.
template <class T> M3<T> basis(uint32_t axis, Q4<T> q) {
    const V3<T> e[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    return M3<T>{{qrotate(q, e[axis]), qrotate(q, e[(axis + 1) % 3]), qrotate(q, e[(axis + 2) % 3])}};
}

// ---------------------------------------------------------------------------
// Scene view in scalar type T.
// ---------------------------------------------------------------------------
template <class T>
struct Scene {
    const vxo_scene *s;
    V3<T> pos(uint32_t i) const {
        return {T(s->positions[3 * i]), T(s->positions[3 * i + 1]), T(s->positions[3 * i + 2])};
    }
    Q4<T> rot(uint32_t i) const {
        const double *q = s->rotations + 4 * i;
        return {T(q[0]), T(q[1]), T(q[2]), T(q[3])};
    }
    V3<T> fib(uint32_t i) const {
        return {T(s->fibers[3 * i]), T(s->fibers[3 * i + 1]), T(s->fibers[3 * i + 2])};
    }
    V3<T> bond3(const double *a, uint32_t axis, uint32_t i) const {
        const double *p = a + 3 * (size_t(axis) * s->n + i);
        return {T(p[0]), T(p[1]), T(p[2])};
    }
};

// tensor_compliance (src/solver.rs:37-63).
template <class T> M3<T> tensor_compliance(const Scene<T> &sc, uint32_t n, const M3<T> &frame) {
    const vxo_scene *s = sc.s;
    T solid = std::min(std::max(T(s->solid[n]), T(1e-4)), T(1));
    T young = T(s->young[n]), young_perp = T(s->young_perp[n]), shear = T(s->shear[n]);
    T aniso = T(s->anisotropy[n]);
    V3<T> fiber = transpose(frame) * qrotate(sc.rot(n), sc.fib(n));
    T transverse = young + aniso * (young_perp - young);
    T normal = transverse + (young - transverse) * fiber.x * fiber.x;
    V3<T> diagonal = V3<T>{normal, shear, shear} * solid;
    T mn = std::min(std::min(diagonal.x, diagonal.y), diagonal.z);
    if (!finite3(diagonal) || mn <= T(0)) return M3<T>::zero();
    V3<T> inv_diag{T(1) / diagonal.x, T(1) / diagonal.y, T(1) / diagonal.z};
    V3<T> sf = fiber * inv_diag;
    T alpha = shear * aniso * T(0.2);
    M3<T> rank_one{{sf * sf.x, sf * sf.y, sf * sf.z}};
    return M3<T>::diag(inv_diag) - rank_one * (alpha / (T(1) + alpha * dot(fiber, sf)));
}

// stiffness_matrix (src/solver.rs:65-73): series (harmonic) combination of the
// two half-bond compliances, times 2*AREA/CELL.
template <class T> M3<T> stiffness_matrix(const Scene<T> &sc, uint32_t i, uint32_t j, const M3<T> &frame) {
    M3<T> a = tensor_compliance(sc, i, frame);
    M3<T> b = tensor_compliance(sc, j, frame);
    if (mat_is_zero(a) || mat_is_zero(b)) return M3<T>::zero();
    const T cell = T(kCell);
    const T area = cell * cell;
    return inverse(a + b) * (T(2) * area / cell);
}

// bond_thermal_rest (src/solver.rs:78-95).
template <class T> V3<T> bond_thermal_rest(const Scene<T> &sc, uint32_t i, uint32_t j, const M3<T> &frame) {
    const vxo_scene *s = sc.s;
    V3<T> direction = normalize_or_zero(frame.c[0]);
    auto endpoint = [&](uint32_t n) {
        V3<T> fiber = normalize_or_zero(qrotate(sc.rot(n), sc.fib(n)));
        T ap = T(s->alpha_par[n]), aq = T(s->alpha_perp[n]);
        V3<T> ad = direction * aq + fiber * ((ap - aq) * dot(fiber, direction));
        return ad * T(s->delta_temp[n]);
    };
    const T cell = T(kCell);
    return transpose(frame) * (direction * cell + (endpoint(i) + endpoint(j)) * (T(0.5) * cell));
}

template <class T>
struct Constraint {
    uint32_t i, j, axis;
    M3<T> frame, material_k, k;
    V3<T> rest;
};

// make_constraint (src/solver.rs:1325-1345), plus the per-bond stiffness scale
// applied to material_k before the damage factor.
template <class T> Constraint<T> make_constraint(const Scene<T> &sc, uint32_t i, uint32_t axis, uint32_t j) {
    const vxo_scene *s = sc.s;
    size_t b = size_t(axis) * s->n + i;
    Constraint<T> c;
    c.i = i;
    c.j = j;
    c.axis = axis;
    c.frame = basis(axis, slerp_half(sc.rot(i), sc.rot(j)));
    c.material_k = stiffness_matrix(sc, i, j, c.frame);
    if (s->kscale) c.material_k = c.material_k * T(s->kscale[b]);
    T dmg = s->damage ? T(s->damage[b]) : T(0);
    M3<T> local = c.material_k * std::max(T(1) - dmg, T(0.001));
    c.k = c.frame * local * transpose(c.frame);
    V3<T> thermal = bond_thermal_rest(sc, i, j, c.frame);
    V3<T> off = s->formation_offset ? sc.bond3(s->formation_offset, axis, i) : V3<T>{0, 0, 0};
    V3<T> pl = s->plastic_strain ? sc.bond3(s->plastic_strain, axis, i) : V3<T>{0, 0, 0};
    c.rest = c.frame * (thermal + off + pl * T(kCell));
    return c;
}

// ---------------------------------------------------------------------------
// Fracture law (src/fracture.rs) in scalar type T; energies in double.
// ---------------------------------------------------------------------------
template <class T> struct Limits {
    T tension, compression, shear, yield, hardening, plastic_failure, ductility, gc;
};

// directional (src/fracture.rs:73-75).
template <class T> T directional(T parallel, T perpendicular, T alignment, T anisotropy) {
    T a = std::min(std::max(std::fabs(alignment), T(0)), T(1));
    return parallel + anisotropy * (T(1) - a) * (perpendicular - parallel);
}

// pair_limits (src/fracture.rs:76-133) for strength rows a, b.
template <class T> Limits<T> pair_limits(const double *a, const double *b, const T (&alignment)[2],
                                          const T (&anisotropy)[2]) {
    auto one = [&](const double *m, int k) {
        Limits<T> l;
        const int mode = int(m[VXO_S_MODE]);
        l.ductility = mode == VXO_BRITTLE ? T(0)
                      : mode == VXO_DUCTILE ? T(1)
                                            : (T(1) - anisotropy[k]) * T(0.5) + anisotropy[k] * (alignment[k] * alignment[k]);
        l.tension = directional(T(m[VXO_S_TENSION]), T(m[VXO_S_TENSION_PERP]), alignment[k], anisotropy[k]);
        l.compression = directional(T(m[VXO_S_COMPRESSION]), T(m[VXO_S_COMPRESSION_PERP]), alignment[k], anisotropy[k]);
        l.shear = directional(T(m[VXO_S_SHEAR]), T(m[VXO_S_SHEAR_PERP]), alignment[k], anisotropy[k]);
        l.yield = T(m[VXO_S_YIELD]);
        l.hardening = T(m[VXO_S_HARDENING]);
        l.plastic_failure = T(m[VXO_S_PLASTIC_FAILURE]);
        l.gc = directional(T(m[VXO_S_GC]), T(m[VXO_S_GC_PERP]), alignment[k], anisotropy[k]);
        return l;
    };
    Limits<T> x = one(a, 0), y = one(b, 1);
    return Limits<T>{std::min(x.tension, y.tension),     std::min(x.compression, y.compression),
                     std::min(x.shear, y.shear),         std::min(x.yield, y.yield),
                     std::min(x.hardening, y.hardening), std::min(x.plastic_failure, y.plastic_failure),
                     std::min(x.ductility, y.ductility), std::min(x.gc, y.gc)};
}

template <class T> struct BondT {
    uint32_t neighbor;
    V3<T> plastic;
    T damage, eq_plastic, plastic_work, fatigue_damage, previous, reversal, increment;
};
template <class T> BondT<T> bond_in(const vxo_bond &b) {
    return BondT<T>{b.neighbor, V3<T>{T(b.plastic_strain[0]), T(b.plastic_strain[1]), T(b.plastic_strain[2])},
                    T(b.damage), T(b.equivalent_plastic_strain), T(b.plastic_work), T(b.fatigue_damage),
                    T(b.previous_stress_ratio), T(b.reversal_stress_ratio), T(b.loading_increment)};
}
template <class T> void bond_out(const BondT<T> &t, vxo_bond &b) {
    b.neighbor = t.neighbor;
    b.plastic_strain[0] = double(t.plastic.x);
    b.plastic_strain[1] = double(t.plastic.y);
    b.plastic_strain[2] = double(t.plastic.z);
    b.damage = double(t.damage);
    b.equivalent_plastic_strain = double(t.eq_plastic);
    b.plastic_work = double(t.plastic_work);
    b.fatigue_damage = double(t.fatigue_damage);
    b.previous_stress_ratio = double(t.previous);
    b.reversal_stress_ratio = double(t.reversal);
    b.loading_increment = double(t.increment);
}

// energy (src/fracture.rs:186-188).
template <class T> double bond_energy(V3<T> u, const M3<T> &k, T damage) {
    return double(std::max(T(0.5) * (T(1) - damage) * dot(u, k * u), T(0)));
}

// fatigue (src/fracture.rs:190-207): only stress reversals consume life.
template <class T> void fatigue(BondT<T> &b, T signed_ratio, const vxo_fracture_config *c) {
    const T delta = signed_ratio - b.previous;
    if (std::fabs(delta) > T(1e-6)) {
        if (b.increment * delta < T(0)) {
            const T amplitude = std::fabs(b.previous - b.reversal) * T(0.5);
            const T threshold = T(c->fatigue_threshold);
            if (amplitude > threshold) {
                const T level = std::min((amplitude - threshold) / (T(1) - threshold), T(1));
                b.fatigue_damage = std::min(
                    b.fatigue_damage + T(0.5) * T(std::pow(level, T(c->fatigue_exponent))) / T(c->fatigue_cycles), T(1));
            }
            b.reversal = b.previous;
        }
        b.increment = delta;
    }
    b.previous = signed_ratio;
}

// update (src/fracture.rs:210-292).
template <class T>
vxo_dissipation fracture_update(BondT<T> &b, V3<T> u_total, const M3<T> &k, const Limits<T> &l, T dt,
                                const vxo_fracture_config *c) {
    vxo_dissipation r{0, 0, 0, 0, 0};
    if (b.neighbor == kEmpty) return r;
    const T cell = T(kCell), area = cell * cell;
    const T old_damage = b.damage;
    V3<T> u = u_total - b.plastic * cell;
    const double before = bond_energy(u, k, old_damage);
    const V3<T> trial = (k * u) / area;
    if (l.ductility > T(1e-5)) {
        const T equivalent = std::sqrt(dot(trial, trial));
        const T yield_now = l.yield + l.hardening * b.eq_plastic;
        if (equivalent > yield_now) {
            const V3<T> direction = trial / equivalent;
            const V3<T> compliance_direction = (inverse(k) * (direction * area)) / cell;
            const T elastic_modulus = T(1) / std::max(std::sqrt(dot(compliance_direction, compliance_direction)), T(1e-20));
            const T dep = (equivalent - yield_now) / (elastic_modulus + l.hardening) * l.ductility;
            const V3<T> dp = normalize_or_zero(compliance_direction) * dep;
            b.plastic = b.plastic + dp;
            b.eq_plastic += dep;
            u = u_total - b.plastic * cell;
            const double lost = std::max(before - bond_energy(u, k, old_damage), 0.0);
            b.plastic_work += T(lost);
            r.plastic_heat = lost * double(T(c->plastic_heat_fraction));
            r.stored_plastic = lost - r.plastic_heat;
        }
    }
    const V3<T> stress = (k * u) / area;
    const T normal_limit = stress.x >= T(0) ? l.tension : l.compression;
    const T sx = stress.x / std::max(normal_limit, T(1));
    const T sh = std::max(l.shear, T(1));
    const T ratio = std::sqrt(sx * sx + (stress.y * stress.y + stress.z * stress.z) / (sh * sh));
    T principal;
    if (std::fabs(stress.x) >= std::max(std::fabs(stress.y), std::fabs(stress.z)))
        principal = stress.x;
    else if (std::fabs(stress.y) >= std::fabs(stress.z))
        principal = stress.y;
    else
        principal = stress.z;
    fatigue(b, std::copysign(ratio, principal), c);
    if (ratio >= T(1)) {
        if (l.ductility <= T(1e-5))
            b.damage = T(1);
        else
            b.damage = std::min(b.damage + T(c->damage_rate) * dt * std::min(ratio, T(10)), T(1));
    }
    if (l.ductility > T(1e-5) && b.eq_plastic >= std::max(l.plastic_failure, T(1e-6)))
        b.damage = std::min(b.damage + T(c->damage_rate) * dt, T(1));
    if (b.fatigue_damage >= T(1)) b.damage = T(1);
    const double lost = std::max(bond_energy(u, k, old_damage) - bond_energy(u, k, b.damage), 0.0);
    const double budget = double(l.gc) * double(area) * double(b.damage - old_damage);
    r.surface_work = std::min(lost, budget);
    const double available = lost - r.surface_work;
    r.released = available * double(T(c->energy_release));
    r.damage_heat = available - r.released;
    if (b.damage >= T(1)) b.neighbor = kEmpty;
    return r;
}

// fit_rotations (src/solver.rs:327-417).
template <class T> T fit_rotations(vxo_scene *s) {
    Scene<T> sc{s};
    const uint32_t n = s->n;
    std::vector<M3<T>> grad(n, M3<T>::zero());
    std::vector<V3<T>> counts(n, V3<T>{0, 0, 0});
    for (uint32_t axis = 0; axis < 3; ++axis) {
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t j = s->neighbor[size_t(axis) * n + i];
            if (j == kEmpty) continue;
            V3<T> d = (sc.pos(j) - sc.pos(i)) / T(kCell);
            if (!s->kinematic[i]) {
                grad[i].c[axis] = grad[i].c[axis] + d;
                counts[i][axis] += T(1);
            }
            if (!s->kinematic[j]) {
                grad[j].c[axis] = grad[j].c[axis] + d;
                counts[j][axis] += T(1);
            }
        }
    }
    T max_angle = T(0);
    for (uint32_t i = 0; i < n; ++i) {
        if (s->kinematic[i]) continue;
        Q4<T> prev = sc.rot(i);
        M3<T> old = mat_from_quat(prev);
        M3<T> f = grad[i];
        for (int a = 0; a < 3; ++a)
            f.c[a] = counts[i][a] > T(0) ? f.c[a] / counts[i][a] : old.c[a];
        T angle = T(0);
        if (determinant(f) > T(1e-6)) {
            for (int it = 0; it < 5; ++it) {
                M3<T> next = (f + transpose(inverse(f))) * T(0.5);
                T delta = std::max(std::max(max_abs_element(next.c[0] - f.c[0]),
                                            max_abs_element(next.c[1] - f.c[1])),
                                   max_abs_element(next.c[2] - f.c[2]));
                f = next;
                if (delta <= T(1e-6)) break;
            }
            V3<T> x = normalize_or_zero(f.c[0]);
            V3<T> y = normalize_or_zero(f.c[1] - x * dot(x, f.c[1]));
            Q4<T> q = prev;
            if (dot(x, x) > T(0.5) && dot(y, y) > T(0.5)) {
                q = qnormalize(quat_from_mat(M3<T>{{x, y, cross(x, y)}}));
                double *o = s->rotations + 4 * size_t(i);
                o[0] = double(q.x);
                o[1] = double(q.y);
                o[2] = double(q.z);
                o[3] = double(q.w);
            }
            T c = std::min(std::max(std::fabs(qdot(prev, q)), T(0)), T(1));
            angle = T(2) * std::acos(c);
        }
        max_angle = std::max(max_angle, angle);
    }
    return max_angle;
}

// ---------------------------------------------------------------------------
// Cholesky helpers (src/solver.rs:1347-1384).
// ---------------------------------------------------------------------------
template <class T> M3<T> cholesky_lower(const M3<T> &m) {
    T l00 = std::sqrt(std::max(m.c[0].x, T(1e-20)));
    T l10 = m.c[0].y / l00;
    T l20 = m.c[0].z / l00;
    T l11 = std::sqrt(std::max(m.c[1].y - l10 * l10, T(1e-20)));
    T l21 = (m.c[1].z - l20 * l10) / l11;
    T l22 = std::sqrt(std::max(m.c[2].z - l20 * l20 - l21 * l21, T(1e-20)));
    return M3<T>{{{l00, l10, l20}, {0, l11, l21}, {0, 0, l22}}};
}
template <class T> V3<T> solve_lower(const M3<T> &l, V3<T> b) {
    T x = b.x / l.c[0].x;
    T y = (b.y - l.c[0].y * x) / l.c[1].y;
    T z = (b.z - l.c[0].z * x - l.c[1].z * y) / l.c[2].z;
    return {x, y, z};
}
template <class T> V3<T> solve_lower_transpose(const M3<T> &l, V3<T> b) {
    T z = b.z / l.c[2].z;
    T y = (b.y - l.c[1].z * z) / l.c[1].y;
    T x = (b.x - l.c[0].y * y - l.c[0].z * z) / l.c[0].x;
    return {x, y, z};
}
template <class T> M3<T> right_solve_lower_transpose(const M3<T> &m, const M3<T> &l) {
    V3<T> r0 = solve_lower(l, V3<T>{m.c[0].x, m.c[1].x, m.c[2].x});
    V3<T> r1 = solve_lower(l, V3<T>{m.c[0].y, m.c[1].y, m.c[2].y});
    V3<T> r2 = solve_lower(l, V3<T>{m.c[0].z, m.c[1].z, m.c[2].z});
    return M3<T>{{{r0.x, r1.x, r2.x}, {r0.y, r1.y, r2.y}, {r0.z, r1.z, r2.z}}};
}

// ---------------------------------------------------------------------------
// Assembled system (the reference's SolverScratch after solve_coupled's
// assembly, src/solver.rs:1552-1672).
// ---------------------------------------------------------------------------
template <class T>
struct System {
    uint32_t n = 0;
    std::vector<V3<T>> free;
    std::vector<uint32_t> active;
    std::vector<uint32_t> active_index;
    std::vector<M3<T>> diagonal;
    std::vector<V3<T>> rhs;
    std::vector<std::array<uint32_t, 6>> incidence;
    std::vector<uint8_t> incidence_len;
    std::vector<Constraint<T>> constraints;
    std::vector<std::array<uint32_t, 6>> row_neighbors;
    std::vector<std::array<M3<T>, 6>> row_blocks;
    // preconditioner factors
    std::vector<M3<T>> factor_matrix, factor_diagonal, factor_edges;
};

template <class T> bool is_zero_free(V3<T> f) { return f.x == T(0) && f.y == T(0) && f.z == T(0); }

template <class T>
void assemble(System<T> &sys, const vxo_scene *s, double h_in, const double *predicted,
              const uint8_t *free_mask, const double *loads) {
    Scene<T> sc{s};
    const uint32_t n = s->n;
    const T h = T(h_in);
    sys.n = n;
    sys.free.assign(n, V3<T>{0, 0, 0});
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) sys.free[i][a] = ((free_mask[i] >> a) & 1u) ? T(1) : T(0);
    sys.active.clear();
    sys.active_index.assign(n, kEmpty);
    for (uint32_t i = 0; i < n; ++i) {
        if (!is_zero_free(sys.free[i])) {
            sys.active_index[i] = uint32_t(sys.active.size());
            sys.active.push_back(i);
        }
    }
    std::vector<T> masses(n);
    for (uint32_t i = 0; i < n; ++i) masses[i] = T(s->mass[i]) / (h * h);
    sys.incidence.assign(n, {kEmpty, kEmpty, kEmpty, kEmpty, kEmpty, kEmpty});
    sys.incidence_len.assign(n, 0);
    sys.diagonal.assign(n, M3<T>::zero());
    sys.rhs.assign(n, V3<T>{0, 0, 0});
    for (uint32_t i = 0; i < n; ++i) {
        sys.diagonal[i] = M3<T>::diag(V3<T>{masses[i], masses[i], masses[i]});
        V3<T> pred{T(predicted[3 * i]), T(predicted[3 * i + 1]), T(predicted[3 * i + 2])};
        V3<T> load = loads ? V3<T>{T(loads[3 * i]), T(loads[3 * i + 1]), T(loads[3 * i + 2])}
                           : V3<T>{0, 0, 0};
        sys.rhs[i] = (pred - sc.pos(i)) * masses[i] + load;
    }
    sys.constraints.clear();
    for (uint32_t axis = 0; axis < 3; ++axis) {
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t j = s->neighbor[size_t(axis) * n + i];
            if (j == kEmpty) continue;
            bool fi = is_zero_free(sys.free[i]), fj = is_zero_free(sys.free[j]);
            if (fi && fj && s->kinematic[i]) continue;
            Constraint<T> c = make_constraint(sc, i, axis, j);
            if (fi && fj) {
                sys.constraints.push_back(c);
                continue;
            }
            V3<T> force = c.k * (sc.pos(j) - sc.pos(i) - c.rest);
            sys.rhs[i] = sys.rhs[i] + force;
            sys.rhs[j] = sys.rhs[j] - force;
            sys.diagonal[i] = sys.diagonal[i] + c.k;
            sys.diagonal[j] = sys.diagonal[j] + c.k;
            uint32_t edge = uint32_t(sys.constraints.size());
            for (uint32_t v : {i, j}) {
                uint8_t slot = sys.incidence_len[v];
                if (slot < 6) sys.incidence[v][slot] = edge;
                sys.incidence_len[v] = uint8_t(slot + 1);
            }
            sys.constraints.push_back(c);
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        sys.rhs[i] = sys.rhs[i] * sys.free[i];
        M3<T> p = M3<T>::diag(sys.free[i]);
        sys.diagonal[i] = p * sys.diagonal[i] * p + M3<T>::identity() - p;
    }
    // PcgRow build (src/solver.rs:1652-1672): slots in incidence order; slots
    // past incidence_len stay EMPTY with zero blocks.
    const uint32_t na = uint32_t(sys.active.size());
    sys.row_neighbors.assign(na, {kEmpty, kEmpty, kEmpty, kEmpty, kEmpty, kEmpty});
    std::array<M3<T>, 6> zero_blocks;
    zero_blocks.fill(M3<T>::zero());
    sys.row_blocks.assign(na, zero_blocks);
    for (uint32_t slot = 0; slot < na; ++slot) {
        uint32_t i = sys.active[slot];
        M3<T> rp = M3<T>::diag(sys.free[i]);
        for (uint32_t e = 0; e < std::min<uint32_t>(sys.incidence_len[i], 6); ++e) {
            const Constraint<T> &c = sys.constraints[sys.incidence[i][e]];
            uint32_t nb = c.i == i ? c.j : c.i;
            uint32_t nbi = sys.active_index[nb];
            sys.row_neighbors[slot][e] = nbi;
            if (nbi != kEmpty) {
                M3<T> cp = M3<T>::diag(sys.free[nb]);
                sys.row_blocks[slot][e] = -rp * c.k * cp;
            }
        }
    }
}

// build_ic0 (src/solver.rs:1390-1445).
template <class T> void build_ic0(System<T> &sys) {
    const uint32_t n = sys.n;
    sys.factor_matrix.assign(n, M3<T>::zero());
    sys.factor_diagonal.assign(n, M3<T>::identity());
    sys.factor_edges.assign(sys.constraints.size(), M3<T>::zero());
    for (uint32_t i = 0; i < n; ++i) {
        M3<T> p = M3<T>::diag(sys.free[i]);
        sys.factor_matrix[i] = p * sys.diagonal[i] * p + M3<T>::identity() - p;
    }
    for (uint32_t row = 0; row < n; ++row) {
        if (is_zero_free(sys.free[row])) {
            sys.factor_diagonal[row] = M3<T>::identity();
            continue;
        }
        M3<T> rp = M3<T>::diag(sys.free[row]);
        const uint32_t rlen = std::min<uint32_t>(sys.incidence_len[row], 6);
        for (uint32_t e = 0; e < rlen; ++e) {
            uint32_t edge = sys.incidence[row][e];
            const Constraint<T> &c = sys.constraints[edge];
            if (c.j != row) continue;
            uint32_t col = c.i;
            M3<T> cp = M3<T>::diag(sys.free[col]);
            M3<T> value = -rp * c.k * cp;
            const uint32_t clen = std::min<uint32_t>(sys.incidence_len[col], 6);
            for (uint32_t e2 = 0; e2 < rlen; ++e2) {
                uint32_t rs = sys.incidence[row][e2];
                const Constraint<T> &r = sys.constraints[rs];
                if (r.j != row || r.i >= col) continue;
                for (uint32_t e3 = 0; e3 < clen; ++e3) {
                    uint32_t cs = sys.incidence[col][e3];
                    const Constraint<T> &q = sys.constraints[cs];
                    if (q.j == col && q.i == r.i) {
                        value = value - sys.factor_edges[rs] * transpose(sys.factor_edges[cs]);
                        break;
                    }
                }
            }
            sys.factor_edges[edge] = right_solve_lower_transpose(value, sys.factor_diagonal[col]);
            sys.factor_matrix[row] =
                sys.factor_matrix[row] - sys.factor_edges[edge] * transpose(sys.factor_edges[edge]);
        }
        sys.factor_diagonal[row] = cholesky_lower(sys.factor_matrix[row]);
    }
}

// build_block_jacobi (src/solver.rs:1447-1453).
template <class T> void build_block_jacobi(System<T> &sys) {
    sys.factor_diagonal.assign(sys.n, M3<T>::identity());
    for (uint32_t i = 0; i < sys.n; ++i) sys.factor_diagonal[i] = cholesky_lower(sys.diagonal[i]);
}

// apply_ic0 (src/solver.rs:1455-1496): sequential forward / backward sweeps.
template <class T>
void apply_ic0(const System<T> &sys, const std::vector<V3<T>> &rhs, std::vector<V3<T>> &work,
               std::vector<V3<T>> &out) {
    const uint32_t na = uint32_t(sys.active.size());
    for (uint32_t slot = 0; slot < na; ++slot) {
        uint32_t i = sys.active[slot];
        V3<T> value = rhs[slot];
        for (uint32_t e = 0; e < std::min<uint32_t>(sys.incidence_len[i], 6); ++e) {
            uint32_t edge = sys.incidence[i][e];
            const Constraint<T> &c = sys.constraints[edge];
            if (c.j == i && sys.active_index[c.i] != kEmpty)
                value = value - sys.factor_edges[edge] * work[sys.active_index[c.i]];
        }
        work[slot] = solve_lower(sys.factor_diagonal[i], value);
    }
    for (uint32_t k = na; k-- > 0;) {
        uint32_t i = sys.active[k];
        V3<T> value = work[k];
        for (uint32_t e = 0; e < std::min<uint32_t>(sys.incidence_len[i], 6); ++e) {
            uint32_t edge = sys.incidence[i][e];
            const Constraint<T> &c = sys.constraints[edge];
            if (c.i == i && sys.active_index[c.j] != kEmpty)
                value = value - transpose(sys.factor_edges[edge]) * out[sys.active_index[c.j]];
        }
        out[k] = solve_lower_transpose(sys.factor_diagonal[i], value);
    }
}

// The reference gates its rayon loops on the row count (src/solver.rs:14-25).
struct Par {
    bool on;
    int threads;
};

template <class T>
void apply_preconditioner(const System<T> &sys, int precond, const std::vector<V3<T>> &rhs,
                          std::vector<V3<T>> &work, std::vector<V3<T>> &out, Par par) {
    if (precond == VXO_PRECOND_IC0) {
        apply_ic0(sys, rhs, work, out);
        return;
    }
    const int64_t na = int64_t(sys.active.size());
#pragma omp parallel for if (par.on) num_threads(par.threads) schedule(static)
    for (int64_t slot = 0; slot < na; ++slot) {
        const M3<T> &l = sys.factor_diagonal[sys.active[slot]];
        out[slot] = solve_lower_transpose(l, solve_lower(l, rhs[slot]));
    }
}

// CPU PCG (src/solver.rs:1673-1911).
template <class T>
void solve_cpu(System<T> &sys, int precond, uint32_t max_iterations, double tolerance, int threads,
               uint32_t cutoff, double *dx_out, uint32_t *iters_out, double *residual_out,
               double *seconds_out) {
    if (precond == VXO_PRECOND_IC0)
        build_ic0(sys);
    else
        build_block_jacobi(sys);
    const uint32_t na = uint32_t(sys.active.size());
    int nthreads = threads;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    Par par{na >= cutoff && nthreads > 1, std::max(nthreads, 1)};
    std::vector<V3<T>> dx(na, V3<T>{0, 0, 0}), ap(na), r(na), z(na), p(na), work(na);
    for (uint32_t slot = 0; slot < na; ++slot) r[slot] = sys.rhs[sys.active[slot]];
    const int64_t nn = na;
    auto apply = [&](const std::vector<V3<T>> &v, std::vector<V3<T>> &out) {
#pragma omp parallel for if (par.on) num_threads(par.threads) schedule(static)
        for (int64_t slot = 0; slot < nn; ++slot) {
            uint32_t i = sys.active[slot];
            V3<T> value = sys.diagonal[i] * v[slot];
            for (uint32_t e = 0; e < std::min<uint32_t>(sys.incidence_len[i], 6); ++e) {
                uint32_t nb = sys.row_neighbors[slot][e];
                if (nb != kEmpty) value = value + sys.row_blocks[slot][e] * v[nb];
            }
            out[slot] = value;
        }
    };
    // Per-row dot in the vector's scalar type, widened and summed in f64.
    auto vdot = [&](const std::vector<V3<T>> &a, const std::vector<V3<T>> &b) {
        double sum = 0.0;
#pragma omp parallel for if (par.on) num_threads(par.threads) reduction(+ : sum) schedule(static)
        for (int64_t k = 0; k < nn; ++k) sum += double(dot(a[k], b[k]));
        return sum;
    };
    auto t0 = std::chrono::steady_clock::now();
    apply_preconditioner(sys, precond, r, work, z, par);
    p = z;
    double rz = vdot(r, z);
    double initial = std::sqrt(vdot(r, r));
    double residual = 0.0;
    uint32_t iterations = 0;
    for (uint32_t it = 0; it < max_iterations; ++it) {
        if (initial <= 1e-12) break;
        apply(p, ap);
        iterations += 1;
        double denom = vdot(p, ap);
        if (denom <= 1e-30) break;
        T alpha = T(rz / denom);
#pragma omp parallel for if (par.on) num_threads(par.threads) schedule(static)
        for (int64_t k = 0; k < nn; ++k) {
            dx[k] = dx[k] + p[k] * alpha;
            r[k] = r[k] - ap[k] * alpha;
        }
        residual = double(T(std::sqrt(vdot(r, r)) / initial));
        if (residual < tolerance) break;
        apply_preconditioner(sys, precond, r, work, z, par);
        double next = vdot(r, z);
        T beta = T(next / rz);
#pragma omp parallel for if (par.on) num_threads(par.threads) schedule(static)
        for (int64_t k = 0; k < nn; ++k) p[k] = z[k] + p[k] * beta;
        rz = next;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds_out) *seconds_out = std::chrono::duration<double>(t1 - t0).count();
    for (uint32_t k = 0; k < na; ++k) {
        dx_out[3 * k] = double(dx[k].x);
        dx_out[3 * k + 1] = double(dx[k].y);
        dx_out[3 * k + 2] = double(dx[k].z);
    }
    if (iters_out) *iters_out = iterations;
    if (residual_out) *residual_out = residual;
}

// ---------------------------------------------------------------------------
// Device PCG semantics over the reference packing (src/cuda_solver.rs:301-501,
// kernels src/cuda/pcg.ptx). In f32 the per-element updates are single fused
// multiply-adds, as in the kernels.
// ---------------------------------------------------------------------------
template <class T> T fma_t(T a, T b, T c) { return std::fma(a, b, c); }

template <class T>
void matvec_packed(uint32_t n, const uint32_t *nbr, const T *diag, const T *blocks, const T *x, T *y) {
    // block_sparse_matvec (pcg.ptx:4-103): row-major diagonal, then each
    // non-empty slot's row-major block, all as fused multiply-adds.
    for (uint32_t i = 0; i < n; ++i) {
        const T *d = diag + 9 * size_t(i);
        const T *xi = x + 3 * size_t(i);
        T acc[3];
        for (int r = 0; r < 3; ++r) {
            T v = d[3 * r] * xi[0];
            v = fma_t(d[3 * r + 1], xi[1], v);
            v = fma_t(d[3 * r + 2], xi[2], v);
            acc[r] = v;
        }
        for (int sl = 0; sl < 6; ++sl) {
            uint32_t j = nbr[6 * size_t(i) + sl];
            if (j == kEmpty) continue;
            const T *xj = x + 3 * size_t(j);
            const T *b = blocks + 54 * size_t(i) + 9 * sl;
            for (int r = 0; r < 3; ++r) {
                acc[r] = fma_t(b[3 * r], xj[0], acc[r]);
                acc[r] = fma_t(b[3 * r + 1], xj[1], acc[r]);
                acc[r] = fma_t(b[3 * r + 2], xj[2], acc[r]);
            }
        }
        for (int r = 0; r < 3; ++r) y[3 * size_t(i) + r] = acc[r];
    }
}

template <class T> void apply_inverse(uint32_t n, const T *inv, const T *r, T *z) {
    for (uint32_t i = 0; i < n; ++i) {
        const T *m = inv + 9 * size_t(i);
        const T *ri = r + 3 * size_t(i);
        for (int row = 0; row < 3; ++row) {
            T v = m[3 * row] * ri[0];
            v = fma_t(m[3 * row + 1], ri[1], v);
            v = fma_t(m[3 * row + 2], ri[2], v);
            z[3 * size_t(i) + row] = v;
        }
    }
}

// dot_product_single: per row f64 products of the widened elements.
template <class T> double dot_rows_f64(uint32_t n, const T *a, const T *b) {
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const T *x = a + 3 * size_t(i);
        const T *y = b + 3 * size_t(i);
        double v = double(x[0]) * double(y[0]);
        v = std::fma(double(x[1]), double(y[1]), v);
        v = std::fma(double(x[2]), double(y[2]), v);
        sum += v;
    }
    return sum;
}

// scalar_ratio / dot_reduce_ratio_single: guarded quotient cast to the
// vector type.
template <class T> T guarded_ratio(double num, double den) { return den > 1e-30 ? T(num / den) : T(0); }

template <class T>
int solve_device_semantics(uint32_t n, const double *rhs_d, const uint32_t *nbr, const double *diag_d,
                           const double *blocks_d, const double *inv_d, uint32_t max_iterations,
                           double tolerance, uint32_t check_every, double *dx_out, uint32_t *iters_out,
                           double *rel_out) {
    if (n == 0) {
        if (iters_out) *iters_out = 0;
        if (rel_out) *rel_out = 0.0;
        return VXO_OK;
    }
    const size_t m = 3 * size_t(n);
    std::vector<T> rhs(m), diag(9 * size_t(n)), blocks(54 * size_t(n)), inv(9 * size_t(n));
    for (size_t k = 0; k < m; ++k) rhs[k] = T(rhs_d[k]);
    for (size_t k = 0; k < 9 * size_t(n); ++k) {
        diag[k] = T(diag_d[k]);
        inv[k] = T(inv_d[k]);
    }
    for (size_t k = 0; k < 54 * size_t(n); ++k) blocks[k] = T(blocks_d[k]);
    double initial_sq = 0.0;
    for (size_t k = 0; k < m; ++k) initial_sq += double(rhs[k]) * double(rhs[k]);
    const double initial = std::sqrt(initial_sq);
    if (initial <= 1e-12) {
        for (size_t k = 0; k < m; ++k) dx_out[k] = 0.0;
        if (iters_out) *iters_out = 0;
        if (rel_out) *rel_out = 0.0;
        return VXO_OK;
    }
    const T tol = T(tolerance);
    std::vector<T> dx(m, T(0)), r(rhs), z(m), p(m, T(0)), ap(m);
    apply_inverse(n, inv.data(), r.data(), z.data());
    double rho = dot_rows_f64(n, r.data(), z.data());
    for (size_t k = 0; k < m; ++k) p[k] = fma_t(T(0), p[k], z[k]);
    uint32_t iterations = 0;
    double rr_host = std::numeric_limits<double>::infinity();
    for (uint32_t it = 0; it < max_iterations; ++it) {
        matvec_packed(n, nbr, diag.data(), blocks.data(), p.data(), ap.data());
        T alpha = guarded_ratio<T>(rho, dot_rows_f64(n, p.data(), ap.data()));
        for (size_t k = 0; k < m; ++k) {
            dx[k] = fma_t(alpha, p[k], dx[k]);
            r[k] = fma_t(-alpha, ap[k], r[k]);
        }
        apply_inverse(n, inv.data(), r.data(), z.data());
        double rho_next = dot_rows_f64(n, r.data(), z.data());
        T beta = guarded_ratio<T>(rho_next, rho);
        for (size_t k = 0; k < m; ++k) p[k] = fma_t(beta, p[k], z[k]);
        rho = rho_next;
        iterations += 1;
        if ((check_every && iterations % check_every == 0) || iterations == max_iterations) {
            rr_host = dot_rows_f64(n, r.data(), r.data());
            if (!std::isfinite(rr_host) || rr_host < 0.0) return VXO_ERR_NONFINITE_RESIDUAL;
            T rel = T(std::sqrt(rr_host) / initial);
            if (rel <= tol) break;
        }
    }
    if (!std::isfinite(rr_host) || rr_host < 0.0) return VXO_ERR_NONFINITE_RESIDUAL;
    T rel = T(std::sqrt(rr_host) / initial);
    if (!std::isfinite(double(rel)) || rel > tol) {
        if (iters_out) *iters_out = iterations;
        if (rel_out) *rel_out = double(rel);
        return VXO_ERR_NOT_CONVERGED;
    }
    for (size_t k = 0; k < m; ++k)
        if (!std::isfinite(double(dx[k]))) return VXO_ERR_NONFINITE_CORRECTION;
    for (size_t k = 0; k < m; ++k) dx_out[k] = double(dx[k]);
    if (iters_out) *iters_out = iterations;
    if (rel_out) *rel_out = double(rel);
    return VXO_OK;
}

// Exact inverse of a row-major 3x3 via Cholesky-free adjugate in double.
void inverse_row_major(const double *a, double *out) {
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[5] * a[6] - a[3] * a[8];
    double c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double inv = 1.0 / det;
    out[0] = c00 * inv;
    out[1] = (a[2] * a[7] - a[1] * a[8]) * inv;
    out[2] = (a[1] * a[5] - a[2] * a[4]) * inv;
    out[3] = c01 * inv;
    out[4] = (a[0] * a[8] - a[2] * a[6]) * inv;
    out[5] = (a[2] * a[3] - a[0] * a[5]) * inv;
    out[6] = c02 * inv;
    out[7] = (a[1] * a[6] - a[0] * a[7]) * inv;
    out[8] = (a[0] * a[4] - a[1] * a[3]) * inv;
}

template <class T> void export_m3(const M3<T> &m, double *out) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[3 * r + c] = double(m.at(r, c));
}

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
// Functions below are declared extern "C" in vx_oracle.h, so their
// definitions keep C linkage; the file-local templates between them do not.
struct vxo_system {
    int mode;
    System<float> f;
    System<double> d;
};


uint32_t vxo_build_topology(uint32_t n, const int32_t *keys, const uint8_t *kinematic, uint32_t *neighbor) {
    // Occupancy lookup: sort voxel indices by key, then binary-search the
    // +axis neighbour key (equivalent to the reference's hash-map lookup).
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    auto key_less = [&](const int32_t *a, const int32_t *b) {
        if (a[0] != b[0]) return a[0] < b[0];
        if (a[1] != b[1]) return a[1] < b[1];
        return a[2] < b[2];
    };
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return key_less(keys + 3 * size_t(a), keys + 3 * size_t(b)); });
    uint32_t bonds = 0;
    for (uint32_t axis = 0; axis < 3; ++axis) {
        for (uint32_t i = 0; i < n; ++i) {
            int32_t want[3] = {keys[3 * size_t(i)], keys[3 * size_t(i) + 1], keys[3 * size_t(i) + 2]};
            want[axis] += 1;
            auto it = std::lower_bound(order.begin(), order.end(), want, [&](uint32_t v, const int32_t *w) {
                return key_less(keys + 3 * size_t(v), w);
            });
            uint32_t out = kEmpty;
            if (it != order.end()) {
                const int32_t *k = keys + 3 * size_t(*it);
                if (k[0] == want[0] && k[1] == want[1] && k[2] == want[2] &&
                    kinematic[*it] == kinematic[i]) {
                    out = *it;
                    bonds += 1;
                }
            }
            neighbor[size_t(axis) * n + i] = out;
        }
    }
    return bonds;
}

void vxo_bond_frame(int mode, const double *qi, const double *qj, uint32_t axis, double *frame) {
    if (mode == VXO_F32) {
        Q4<float> a{float(qi[0]), float(qi[1]), float(qi[2]), float(qi[3])};
        Q4<float> b{float(qj[0]), float(qj[1]), float(qj[2]), float(qj[3])};
        export_m3(basis(axis, slerp_half(a, b)), frame);
    } else {
        Q4<double> a{qi[0], qi[1], qi[2], qi[3]};
        Q4<double> b{qj[0], qj[1], qj[2], qj[3]};
        export_m3(basis(axis, slerp_half(a, b)), frame);
    }
}

template <class T> static M3<T> import_m3(const double *a) {
    M3<T> m;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m.c[c][r] = T(a[3 * r + c]);
    return m;
}

void vxo_stiffness_matrix(int mode, const vxo_scene *s, uint32_t i, uint32_t j, const double *frame, double *k) {
    if (mode == VXO_F32)
        export_m3(stiffness_matrix(Scene<float>{s}, i, j, import_m3<float>(frame)), k);
    else
        export_m3(stiffness_matrix(Scene<double>{s}, i, j, import_m3<double>(frame)), k);
}

void vxo_bond_thermal_rest(int mode, const vxo_scene *s, uint32_t i, uint32_t j, const double *frame,
                           double *rest) {
    if (mode == VXO_F32) {
        V3<float> r = bond_thermal_rest(Scene<float>{s}, i, j, import_m3<float>(frame));
        rest[0] = r.x, rest[1] = r.y, rest[2] = r.z;
    } else {
        V3<double> r = bond_thermal_rest(Scene<double>{s}, i, j, import_m3<double>(frame));
        rest[0] = r.x, rest[1] = r.y, rest[2] = r.z;
    }
}

double vxo_fit_rotations(int mode, vxo_scene *s) {
    if (mode == VXO_F32) {
        // Round the state to f32 first so the fit sees reference values.
        return double(fit_rotations<float>(s));
    }
    return fit_rotations<double>(s);
}

vxo_system *vxo_assemble(int mode, const vxo_scene *s, double h, const double *predicted,
                         const uint8_t *free_mask, const double *loads) {
    vxo_system *sys = new vxo_system();
    sys->mode = mode;
    if (mode == VXO_F32)
        assemble(sys->f, s, h, predicted, free_mask, loads);
    else
        assemble(sys->d, s, h, predicted, free_mask, loads);
    return sys;
}

void vxo_system_free(vxo_system *sys) { delete sys; }

uint32_t vxo_system_active_count(const vxo_system *sys) {
    return uint32_t(sys->mode == VXO_F32 ? sys->f.active.size() : sys->d.active.size());
}
uint32_t vxo_system_constraint_count(const vxo_system *sys) {
    return uint32_t(sys->mode == VXO_F32 ? sys->f.constraints.size() : sys->d.constraints.size());
}

template <class T>
static void export_rows_t(const System<T> &sys, uint32_t *active, uint32_t *neighbors, double *blocks,
                          double *diagonal, double *inverse_out, double *rhs) {
    const uint32_t na = uint32_t(sys.active.size());
    for (uint32_t slot = 0; slot < na; ++slot) {
        uint32_t i = sys.active[slot];
        if (active) active[slot] = i;
        if (neighbors)
            for (int e = 0; e < 6; ++e) neighbors[6 * size_t(slot) + e] = sys.row_neighbors[slot][e];
        if (blocks)
            for (int e = 0; e < 6; ++e) export_m3(sys.row_blocks[slot][e], blocks + 54 * size_t(slot) + 9 * e);
        if (diagonal) export_m3(sys.diagonal[i], diagonal + 9 * size_t(slot));
        // The packed inverse is Mat3::inverse of the projected diagonal
        // (src/solver.rs:1765), evaluated in the system's scalar type.
        if (inverse_out) export_m3(inverse(sys.diagonal[i]), inverse_out + 9 * size_t(slot));
        if (rhs) {
            rhs[3 * size_t(slot)] = double(sys.rhs[i].x);
            rhs[3 * size_t(slot) + 1] = double(sys.rhs[i].y);
            rhs[3 * size_t(slot) + 2] = double(sys.rhs[i].z);
        }
    }
}

void vxo_system_export_rows(const vxo_system *sys, uint32_t *active, uint32_t *neighbors, double *blocks,
                            double *diagonal, double *inverse_out, double *rhs) {
    if (sys->mode == VXO_F32)
        export_rows_t(sys->f, active, neighbors, blocks, diagonal, inverse_out, rhs);
    else
        export_rows_t(sys->d, active, neighbors, blocks, diagonal, inverse_out, rhs);
}

template <class T>
static void export_full_t(const System<T> &sys, double *diagonal, double *rhs, uint32_t *incidence,
                          uint8_t *incidence_len) {
    for (uint32_t i = 0; i < sys.n; ++i) {
        if (diagonal) export_m3(sys.diagonal[i], diagonal + 9 * size_t(i));
        if (rhs) {
            rhs[3 * size_t(i)] = double(sys.rhs[i].x);
            rhs[3 * size_t(i) + 1] = double(sys.rhs[i].y);
            rhs[3 * size_t(i) + 2] = double(sys.rhs[i].z);
        }
        if (incidence)
            for (int e = 0; e < 6; ++e) incidence[6 * size_t(i) + e] = sys.incidence[i][e];
        if (incidence_len) incidence_len[i] = sys.incidence_len[i];
    }
}

void vxo_system_export_full(const vxo_system *sys, double *diagonal, double *rhs, uint32_t *incidence,
                            uint8_t *incidence_len) {
    if (sys->mode == VXO_F32)
        export_full_t(sys->f, diagonal, rhs, incidence, incidence_len);
    else
        export_full_t(sys->d, diagonal, rhs, incidence, incidence_len);
}

template <class T>
static void export_constraints_t(const System<T> &sys, uint32_t *ija, double *frame, double *material_k,
                                 double *k, double *rest) {
    for (size_t e = 0; e < sys.constraints.size(); ++e) {
        const Constraint<T> &c = sys.constraints[e];
        if (ija) {
            ija[3 * e] = c.i;
            ija[3 * e + 1] = c.j;
            ija[3 * e + 2] = c.axis;
        }
        if (frame) export_m3(c.frame, frame + 9 * e);
        if (material_k) export_m3(c.material_k, material_k + 9 * e);
        if (k) export_m3(c.k, k + 9 * e);
        if (rest) {
            rest[3 * e] = double(c.rest.x);
            rest[3 * e + 1] = double(c.rest.y);
            rest[3 * e + 2] = double(c.rest.z);
        }
    }
}

void vxo_system_export_constraints(const vxo_system *sys, uint32_t *ija, double *frame, double *material_k,
                                   double *k, double *rest) {
    if (sys->mode == VXO_F32)
        export_constraints_t(sys->f, ija, frame, material_k, k, rest);
    else
        export_constraints_t(sys->d, ija, frame, material_k, k, rest);
}

void vxo_solve_cpu(vxo_system *sys, int precond, uint32_t max_iterations, double tolerance, int threads,
                   uint32_t parallel_cutoff, double *dx, uint32_t *iterations, double *residual,
                   double *seconds) {
    if (sys->mode == VXO_F32)
        solve_cpu(sys->f, precond, max_iterations, tolerance, threads, parallel_cutoff, dx, iterations,
                  residual, seconds);
    else
        solve_cpu(sys->d, precond, max_iterations, tolerance, threads, parallel_cutoff, dx, iterations,
                  residual, seconds);
}

int vxo_solve_device_semantics(int mode, uint32_t n, const double *rhs, const uint32_t *neighbors,
                               const double *diagonal, const double *blocks, const double *inverse_d,
                               uint32_t max_iterations, double tolerance, uint32_t check_every, double *dx,
                               uint32_t *iterations, double *relative) {
    if (mode == VXO_F32)
        return solve_device_semantics<float>(n, rhs, neighbors, diagonal, blocks, inverse_d, max_iterations,
                                             tolerance, check_every, dx, iterations, relative);
    return solve_device_semantics<double>(n, rhs, neighbors, diagonal, blocks, inverse_d, max_iterations,
                                          tolerance, check_every, dx, iterations, relative);
}

void vxo_matvec(int mode, uint32_t n, const uint32_t *neighbors, const double *diagonal, const double *blocks,
                const double *x, double *y) {
    if (mode == VXO_F32) {
        std::vector<float> d(diagonal, diagonal + 9 * size_t(n)), b(blocks, blocks + 54 * size_t(n)),
            xv(x, x + 3 * size_t(n)), yv(3 * size_t(n));
        matvec_packed<float>(n, neighbors, d.data(), b.data(), xv.data(), yv.data());
        for (size_t k = 0; k < 3 * size_t(n); ++k) y[k] = yv[k];
    } else {
        matvec_packed<double>(n, neighbors, diagonal, blocks, x, y);
    }
}

// Rows per chunk of the exact solve's parallel loops; dot products sum the
// chunks' partial sums in chunk order, so the result does not depend on the
// thread count.
constexpr size_t kExactChunk = 4096;

// y = A x over the reference packing (the same arithmetic as matvec_packed),
// rows split over OpenMP threads.
void matvec_packed_par(uint32_t n, const uint32_t *nbr, const double *diag, const double *blocks, const double *x,
                       double *y) {
    const int64_t chunks = int64_t((size_t(n) + kExactChunk - 1) / kExactChunk);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < chunks; ++c) {
        const size_t r0 = size_t(c) * kExactChunk;
        const uint32_t rows = uint32_t(std::min(kExactChunk, size_t(n) - r0));
        // Row i of the chunk reads neighbours anywhere in x: offset the row
        // arrays only.
        const uint32_t *nb = nbr + 6 * r0;
        const double *d = diag + 9 * r0;
        const double *b = blocks + 54 * r0;
        for (uint32_t i = 0; i < rows; ++i) {
            const double *di = d + 9 * size_t(i);
            const double *xi = x + 3 * (r0 + i);
            double acc[3];
            for (int r = 0; r < 3; ++r) {
                double v = di[3 * r] * xi[0];
                v = fma_t(di[3 * r + 1], xi[1], v);
                v = fma_t(di[3 * r + 2], xi[2], v);
                acc[r] = v;
            }
            for (int sl = 0; sl < 6; ++sl) {
                uint32_t j = nb[6 * size_t(i) + sl];
                if (j == kEmpty) continue;
                const double *xj = x + 3 * size_t(j);
                const double *bl = b + 54 * size_t(i) + 9 * sl;
                for (int r = 0; r < 3; ++r) {
                    acc[r] = fma_t(bl[3 * r], xj[0], acc[r]);
                    acc[r] = fma_t(bl[3 * r + 1], xj[1], acc[r]);
                    acc[r] = fma_t(bl[3 * r + 2], xj[2], acc[r]);
                }
            }
            for (int r = 0; r < 3; ++r) y[3 * (r0 + i) + r] = acc[r];
        }
    }
}

// FP64 block-Jacobi CG to a tight relative residual: the accuracy reference
// ("FP64 oracle") every device solve is compared with. Not a restatement of a
// reference function (the reference has no FP64 solve); textbook PCG with a
// true-residual restart so the returned figure is the residual of x.
double vxo_solve_exact(uint32_t n, const double *rhs, const uint32_t *nbr, const double *diagonal,
                       const double *blocks, double tolerance, uint32_t max_iterations, double *x,
                       uint32_t *iterations_out) {
    const size_t m = 3 * size_t(n);
    const int64_t chunks = int64_t((size_t(n) + kExactChunk - 1) / kExactChunk);
    std::vector<double> inv(9 * size_t(n));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) inverse_row_major(diagonal + 9 * size_t(i), inv.data() + 9 * size_t(i));
    std::vector<double> partial(size_t(std::max<int64_t>(chunks, 1)));
    // Sum over rows of f(row), fixed chunks, partials added in chunk order.
    auto chunked_sum = [&](auto &&f) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < chunks; ++c) {
            const size_t k0 = 3 * size_t(c) * kExactChunk, k1 = std::min(m, k0 + 3 * kExactChunk);
            double s = 0.0;
            for (size_t k = k0; k < k1; ++k) s += f(k);
            partial[size_t(c)] = s;
        }
        double s = 0.0;
        for (int64_t c = 0; c < chunks; ++c) s += partial[size_t(c)];
        return s;
    };
    const double bnorm = std::sqrt(chunked_sum([&](size_t k) { return rhs[k] * rhs[k]; }));
    std::vector<double> r(rhs, rhs + m), z(m), p(m), ap(m);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < int64_t(m); ++k) x[k] = 0.0;
    if (iterations_out) *iterations_out = 0;
    if (bnorm == 0.0) return 0.0;
    auto dotv = [&](const std::vector<double> &a, const std::vector<double> &b) {
        return chunked_sum([&](size_t k) { return a[k] * b[k]; });
    };
    auto precond = [&](const std::vector<double> &in, std::vector<double> &out) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(n); ++i) {
            const double *mi = inv.data() + 9 * size_t(i);
            for (int row = 0; row < 3; ++row)
                out[3 * i + row] = mi[3 * row] * in[3 * i] + mi[3 * row + 1] * in[3 * i + 1] +
                                   mi[3 * row + 2] * in[3 * i + 2];
        }
    };
    auto true_rel = [&]() {
        matvec_packed_par(n, nbr, diagonal, blocks, x, ap.data());
        const double s = chunked_sum([&](size_t k) {
            const double d = rhs[k] - ap[k];
            return d * d;
        });
        return std::sqrt(s) / bnorm;
    };
    uint32_t iterations = 0;
    double rel = 1.0;
    // Restarted CG with a true-residual recompute at each restart so the
    // final figure is the residual of the returned x, not the recurrence.
    while (iterations < max_iterations) {
        precond(r, z);
        p = z;
        double rz = dotv(r, z);
        uint32_t since = 0;
        while (iterations < max_iterations && since < 5000) {
            matvec_packed_par(n, nbr, diagonal, blocks, p.data(), ap.data());
            double pap = dotv(p, ap);
            if (!(pap > 0.0)) break;
            double alpha = rz / pap;
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < int64_t(m); ++k) {
                x[k] += alpha * p[k];
                r[k] -= alpha * ap[k];
            }
            iterations += 1;
            since += 1;
            if (std::sqrt(dotv(r, r)) / bnorm <= tolerance * 0.5) break;
            precond(r, z);
            double rz2 = dotv(r, z);
            double beta = rz2 / rz;
            rz = rz2;
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < int64_t(m); ++k) p[k] = z[k] + beta * p[k];
        }
        rel = true_rel();
        if (rel <= tolerance || since == 0) break;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < int64_t(m); ++k) r[k] = rhs[k] - ap[k];
    }
    if (iterations_out) *iterations_out = iterations;
    return rel;
}

// CG on the red-black reduced system, FP64, over the reference packing.
// Rows are coloured (black[i] = 1 for odd x + y + z); every bond joins a red
// and a black row, so A = [D_rr A_rb; A_br D_bb] with block-diagonal D_rr,
// D_bb. Eliminating the red rows: S x_b = b_b - A_br D_rr^-1 b_r with
// S = D_bb - A_br D_rr^-1 A_rb, solved by CG with block-Jacobi (D_bb^-1);
// then x_r = D_rr^-1 (b_r - A_rb x_b). The loop checks, before each
// iteration, r.r <= tol^2 b.b (b the full right-hand side) or
// r.r <= inner_tol^2 r0.r0 (r0 the reduced residual at x_b = 0). Returns the
// full system's true relative residual of the returned x. history (nullable,
// max_iterations + 1 entries): ||r|| of the reduced system before each
// iteration. No counterpart in the reference (its device preconditioner is
// block-Jacobi, src/cuda_solver.rs:366-500); the algorithm is textbook
// red-black reduction (equivalently CG preconditioned by red-black block
// symmetric Gauss-Seidel in Eisenstat's form).
double vxo_solve_red_black(uint32_t n, const double *rhs, const uint32_t *nbr, const double *diagonal,
                           const double *blocks, const uint8_t *black, double tolerance, double inner_tolerance,
                           uint32_t max_iterations, double *x, uint32_t *iterations_out, double *history) {
    const size_t m = 3 * size_t(n);
    std::vector<double> inv(9 * size_t(n));
    for (uint32_t i = 0; i < n; ++i) inverse_row_major(diagonal + 9 * size_t(i), inv.data() + 9 * size_t(i));
    auto mul3 = [](const double *a, const double *v, double *out) {
        for (int r = 0; r < 3; ++r) out[r] = a[3 * r] * v[0] + a[3 * r + 1] * v[1] + a[3 * r + 2] * v[2];
    };
    // out_i = sum_e B_ie v_j over rows of one colour (the off-diagonal part).
    auto offdiag = [&](const std::vector<double> &v, std::vector<double> &out, uint8_t colour) {
        for (uint32_t i = 0; i < n; ++i) {
            if (black[i] != colour) continue;
            double acc[3] = {0, 0, 0};
            for (int e = 0; e < 6; ++e) {
                const uint32_t j = nbr[6 * size_t(i) + e];
                if (j == kEmpty) continue;
                double t[3];
                mul3(blocks + 54 * size_t(i) + 9 * e, v.data() + 3 * size_t(j), t);
                for (int r = 0; r < 3; ++r) acc[r] += t[r];
            }
            for (int r = 0; r < 3; ++r) out[3 * size_t(i) + r] = acc[r];
        }
    };
    // out_i = M_i v_i over rows of one colour.
    auto blockwise = [&](const double *mats, const std::vector<double> &v, std::vector<double> &out, uint8_t colour) {
        for (uint32_t i = 0; i < n; ++i)
            if (black[i] == colour) mul3(mats + 9 * size_t(i), v.data() + 3 * size_t(i), out.data() + 3 * size_t(i));
    };
    auto dotb = [&](const std::vector<double> &a, const std::vector<double> &b) {
        double s = 0.0;
        for (uint32_t i = 0; i < n; ++i)
            if (black[i])
                for (int r = 0; r < 3; ++r) s += a[3 * size_t(i) + r] * b[3 * size_t(i) + r];
        return s;
    };
    std::vector<double> b(rhs, rhs + m), t(m, 0.0), u(m, 0.0), xs(m, 0.0);
    double bb = 0.0;
    for (size_t k = 0; k < m; ++k) bb += b[k] * b[k];
    // S p on the black rows: D_bb p - A_br D_rr^-1 A_rb p.
    auto apply_s = [&](const std::vector<double> &p, std::vector<double> &out) {
        offdiag(p, t, 0);
        blockwise(inv.data(), t, u, 0);
        offdiag(u, t, 1);
        blockwise(diagonal, p, out, 1);
        for (uint32_t i = 0; i < n; ++i)
            if (black[i])
                for (int r = 0; r < 3; ++r) out[3 * size_t(i) + r] -= t[3 * size_t(i) + r];
    };
    // reduced right-hand side b_b - A_br D_rr^-1 b_r
    std::vector<double> r(m, 0.0), z(m, 0.0), p(m, 0.0), ap(m, 0.0);
    blockwise(inv.data(), b, u, 0);
    offdiag(u, t, 1);
    for (uint32_t i = 0; i < n; ++i)
        if (black[i])
            for (int q = 0; q < 3; ++q) r[3 * size_t(i) + q] = b[3 * size_t(i) + q] - t[3 * size_t(i) + q];
    const double rr0 = dotb(r, r);
    blockwise(inv.data(), r, z, 1);
    double rz = dotb(r, z);
    p = z;
    uint32_t k = 0;
    for (;; ++k) {
        const double rr = dotb(r, r);
        if (history) history[k] = std::sqrt(rr);
        if (!(rr > tolerance * tolerance * bb) || !(rr > inner_tolerance * inner_tolerance * rr0) || k >= max_iterations)
            break;
        if (k > 0) {
            blockwise(inv.data(), r, z, 1);
            const double rz2 = dotb(r, z);
            const double beta = rz2 / rz;
            rz = rz2;
            for (uint32_t i = 0; i < n; ++i)
                if (black[i])
                    for (int q = 0; q < 3; ++q) p[3 * size_t(i) + q] = z[3 * size_t(i) + q] + beta * p[3 * size_t(i) + q];
        }
        apply_s(p, ap);
        const double alpha = rz / dotb(p, ap);
        for (uint32_t i = 0; i < n; ++i)
            if (black[i])
                for (int q = 0; q < 3; ++q) {
                    xs[3 * size_t(i) + q] += alpha * p[3 * size_t(i) + q];
                    r[3 * size_t(i) + q] -= alpha * ap[3 * size_t(i) + q];
                }
    }
    // x_r = D_rr^-1 (b_r - A_rb x_b)
    offdiag(xs, t, 0);
    for (uint32_t i = 0; i < n; ++i)
        if (!black[i])
            for (int q = 0; q < 3; ++q) u[3 * size_t(i) + q] = b[3 * size_t(i) + q] - t[3 * size_t(i) + q];
    blockwise(inv.data(), u, xs, 0);
    for (size_t q = 0; q < m; ++q) x[q] = xs[q];
    if (iterations_out) *iterations_out = k;
    matvec_packed<double>(n, nbr, diagonal, blocks, x, ap.data());
    double s2 = 0.0;
    for (size_t q = 0; q < m; ++q) s2 += (rhs[q] - ap[q]) * (rhs[q] - ap[q]);
    return bb > 0.0 ? std::sqrt(s2 / bb) : 0.0;
}

double vxo_dot_rows(uint32_t n, const double *a, const double *b) {
    std::vector<float> x(a, a + 3 * size_t(n)), y(b, b + 3 * size_t(n));
    return dot_rows_f64<float>(n, x.data(), y.data());
}

double vxo_guarded_ratio(int mode, double num, double den) {
    return mode == VXO_F32 ? double(guarded_ratio<float>(num, den)) : guarded_ratio<double>(num, den);
}

void vxo_pair_limits(int mode, const double *a, const double *b, const double alignment[2], const double anisotropy[2],
                     double limits[VXO_LIMITS]) {
    auto put = [&](auto l) {
        const double v[VXO_LIMITS] = {double(l.tension), double(l.compression), double(l.shear), double(l.yield),
                                      double(l.hardening), double(l.plastic_failure), double(l.ductility), double(l.gc)};
        std::memcpy(limits, v, sizeof v);
    };
    if (mode == VXO_F32) {
        const float al[2] = {float(alignment[0]), float(alignment[1])}, an[2] = {float(anisotropy[0]), float(anisotropy[1])};
        put(pair_limits<float>(a, b, al, an));
    } else {
        const double al[2] = {alignment[0], alignment[1]}, an[2] = {anisotropy[0], anisotropy[1]};
        put(pair_limits<double>(a, b, al, an));
    }
}

template <class T> static void fatigue_t(vxo_bond *bond, double signed_ratio, const vxo_fracture_config *c) {
    BondT<T> b = bond_in<T>(*bond);
    fatigue(b, T(signed_ratio), c);
    bond_out(b, *bond);
}
void vxo_fatigue(int mode, vxo_bond *bond, double signed_ratio, const vxo_fracture_config *c) {
    if (mode == VXO_F32)
        fatigue_t<float>(bond, signed_ratio, c);
    else
        fatigue_t<double>(bond, signed_ratio, c);
}

template <class T>
static void fracture_update_t(vxo_bond *bond, const double u_total[3], const double k[9], const double lim[VXO_LIMITS],
                              double dt, const vxo_fracture_config *c, vxo_dissipation *out) {
    BondT<T> b = bond_in<T>(*bond);
    M3<T> km;
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col) km.c[col][r] = T(k[3 * r + col]);
    const Limits<T> l{T(lim[0]), T(lim[1]), T(lim[2]), T(lim[3]), T(lim[4]), T(lim[5]), T(lim[6]), T(lim[7])};
    vxo_dissipation d = fracture_update(b, V3<T>{T(u_total[0]), T(u_total[1]), T(u_total[2])}, km, l, T(dt), c);
    bond_out(b, *bond);
    if (out) *out = d;
}
void vxo_fracture_update(int mode, vxo_bond *bond, const double u_total[3], const double k[9],
                         const double limits[VXO_LIMITS], double dt, const vxo_fracture_config *c, vxo_dissipation *out) {
    if (mode == VXO_F32)
        fracture_update_t<float>(bond, u_total, k, limits, dt, c, out);
    else
        fracture_update_t<double>(bond, u_total, k, limits, dt, c, out);
}

// The fracture pass of a substep (src/solver.rs:627-665): every constraint of
// the last assembly whose bond is intact and whose -end REV is not kinematic,
// with bond_limits (src/solver.rs:124-161; thermal strength scale 1) and the
// elastic displacement in the constraint's frame (src/solver.rs:106-110).
template <class T>
static void fracture_pass(const System<T> &sys, vxo_scene *s, vxo_fracture_state *fr, T h) {
    Scene<T> sc{s};
    const uint32_t n = s->n;
    for (const Constraint<T> &c : sys.constraints) {
        const uint32_t i = c.i, j = c.j;
        const size_t b = size_t(c.axis) * n + i;
        if (fr->neighbor[b] == kEmpty || s->kinematic[i]) continue;
        const V3<T> direction = c.frame.c[0];
        T alignment[2], anisotropy[2];
        const uint32_t ends[2] = {i, j};
        T solid = T(1);
        for (int e = 0; e < 2; ++e) {
            alignment[e] = dot(direction, qrotate(sc.rot(ends[e]), sc.fib(ends[e])));
            anisotropy[e] = T(s->anisotropy[ends[e]]);
            solid = std::min(solid, std::min(std::max(T(s->solid[ends[e]]), T(0)), T(1)));
        }
        Limits<T> l = pair_limits<T>(fr->strength + size_t(VXO_STRENGTH) * i, fr->strength + size_t(VXO_STRENGTH) * j,
                                     alignment, anisotropy);
        const T scale = solid * solid;
        l.tension *= scale;
        l.compression *= scale;
        l.shear *= scale;
        l.yield *= scale;
        l.hardening *= scale;
        l.gc *= scale;
        const V3<T> off = s->formation_offset ? sc.bond3(s->formation_offset, c.axis, i) : V3<T>{0, 0, 0};
        const V3<T> u = transpose(c.frame) * (sc.pos(j) - sc.pos(i)) - bond_thermal_rest(sc, i, j, c.frame) - off;
        vxo_bond bond;
        bond.neighbor = fr->neighbor[b];
        for (int a = 0; a < 3; ++a) bond.plastic_strain[a] = fr->plastic_strain[3 * b + a];
        bond.damage = fr->damage[b];
        bond.equivalent_plastic_strain = fr->equivalent_plastic_strain[b];
        bond.plastic_work = fr->plastic_work[b];
        bond.fatigue_damage = fr->fatigue_damage[b];
        bond.previous_stress_ratio = fr->previous_stress_ratio[b];
        bond.reversal_stress_ratio = fr->reversal_stress_ratio[b];
        bond.loading_increment = fr->loading_increment[b];
        BondT<T> bt = bond_in<T>(bond);
        const vxo_dissipation d = fracture_update(bt, u, c.material_k, l, h, &fr->config);
        bond_out(bt, bond);
        fr->neighbor[b] = bond.neighbor;
        for (int a = 0; a < 3; ++a) fr->plastic_strain[3 * b + a] = bond.plastic_strain[a];
        fr->damage[b] = bond.damage;
        fr->equivalent_plastic_strain[b] = bond.equivalent_plastic_strain;
        fr->plastic_work[b] = bond.plastic_work;
        fr->fatigue_damage[b] = bond.fatigue_damage;
        fr->previous_stress_ratio[b] = bond.previous_stress_ratio;
        fr->reversal_stress_ratio[b] = bond.reversal_stress_ratio;
        fr->loading_increment[b] = bond.loading_increment;
        fr->dissipated.plastic_heat += d.plastic_heat;
        fr->dissipated.stored_plastic += d.stored_plastic;
        fr->dissipated.surface_work += d.surface_work;
        fr->dissipated.damage_heat += d.damage_heat;
        fr->dissipated.released += d.released;
        if (bond.neighbor == kEmpty) fr->broken += 1;
    }
}

// Per constraint of the last assembly at the solved positions, before the
// fracture pass: the multiplier lambda = -F^T k (p_j - p_i - rest) h^2
// (src/solver.rs:1824-1827, 1921-1927) and, for a non-kinematic first REV, the
// bond force F (K (u - plastic CELL)) (1 - damage) (src/solver.rs:627-640).
// Outputs are (3, n, 3) [axis][voxel][component], zero without a constraint.
template <class T>
static void bond_forces_t(const System<T> &sys, vxo_scene *s, T h, double *last_force, double *lambda) {
    const uint32_t n = s->n;
    Scene<T> sc{s};
    if (last_force) std::fill(last_force, last_force + 9 * size_t(n), 0.0);
    if (lambda) std::fill(lambda, lambda + 9 * size_t(n), 0.0);
    for (const Constraint<T> &c : sys.constraints) {
        const size_t b = size_t(c.axis) * n + c.i;
        if (lambda) {
            V3<T> force = c.k * (sc.pos(c.j) - sc.pos(c.i) - c.rest);
            V3<T> l = transpose(c.frame) * force * (-(h * h));
            lambda[3 * b] = double(l.x);
            lambda[3 * b + 1] = double(l.y);
            lambda[3 * b + 2] = double(l.z);
        }
        if (!last_force || s->kinematic[c.i]) continue;
        V3<T> off = s->formation_offset ? sc.bond3(s->formation_offset, c.axis, c.i) : V3<T>{0, 0, 0};
        V3<T> pl = s->plastic_strain ? sc.bond3(s->plastic_strain, c.axis, c.i) : V3<T>{0, 0, 0};
        V3<T> u = transpose(c.frame) * (sc.pos(c.j) - sc.pos(c.i)) - bond_thermal_rest(sc, c.i, c.j, c.frame) - off;
        T dmg = s->damage ? T(s->damage[b]) : T(0);
        V3<T> f = c.frame * (c.material_k * (u - pl * T(kCell))) * (T(1) - dmg);
        last_force[3 * b] = double(f.x);
        last_force[3 * b + 1] = double(f.y);
        last_force[3 * b + 2] = double(f.z);
    }
}

// Solve part of step_with_axis_targets_and_loads (src/solver.rs:446-570,
// 668-683): predict, impose targets, fit rotations, corotation passes of
// assemble + PCG + apply, then velocities.
template <class T>
static void step_t(vxo_scene *s, const vxo_step_config *cfg, const double *targets, const double *loads,
                   vxo_step_report *report, double *last_force, double *lambda, vxo_fracture_state *fracture = nullptr,
                   vxo_tool *tool = nullptr) {
    const uint32_t n = s->n;
    const T h = T(cfg->fixed_dt) / T(cfg->substeps);
    const V3<T> g{T(cfg->gravity[0]), T(cfg->gravity[1]), T(cfg->gravity[2])};
    auto has_target = [&](uint32_t i, int a) { return targets && !std::isnan(targets[3 * size_t(i) + a]); };
    std::vector<uint8_t> free_mask(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint8_t m = 0;
        for (int a = 0; a < 3; ++a)
            if (!s->kinematic[i] && !has_target(i, a)) m |= uint8_t(1u << a);
        free_mask[i] = m;
    }
    std::vector<double> applied(3 * size_t(n), 0.0);
    for (uint32_t i = 0; i < n; ++i)
        if (!s->kinematic[i] && loads)
            for (int a = 0; a < 3; ++a) applied[3 * i + a] = double(T(loads[3 * i + a]));
    vxo_step_report rep{0, 0, 0.0, 0.0};
    System<T> sys;
    for (uint32_t sub = 0; sub < cfg->substeps; ++sub) {
        std::vector<double> previous(s->positions, s->positions + 3 * size_t(n));
        std::vector<V3<T>> drift(n, V3<T>{0, 0, 0});
        for (uint32_t i = 0; i < n; ++i) {
            if (s->kinematic[i]) continue;
            V3<T> v{T(s->velocities[3 * i]), T(s->velocities[3 * i + 1]), T(s->velocities[3 * i + 2])};
            v = v + g * h;
            for (int a = 0; a < 3; ++a)
                if (has_target(i, a)) v[a] = T(0);
            drift[i] = v;
            V3<T> p{T(s->positions[3 * i]), T(s->positions[3 * i + 1]), T(s->positions[3 * i + 2])};
            p = p + v * h;
            for (int a = 0; a < 3; ++a) {
                s->velocities[3 * i + a] = double(v[a]);
                s->positions[3 * i + a] = double(p[a]);
            }
        }
        for (uint32_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a)
                if (has_target(i, a)) s->positions[3 * i + a] = double(T(targets[3 * i + a]));
        // The tool's contact impulses, acting near the middle of the substep
        // (src/solver.rs:530-543).
        std::vector<double> contact(3 * size_t(n), 0.0);
        const double gd[3] = {cfg->gravity[0], cfg->gravity[1], cfg->gravity[2]};
        std::vector<double> mass;
        if (tool) {
            mass.assign(s->mass, s->mass + n);
            vxo_tool_advance(std::is_same<T, float>::value ? VXO_F32 : VXO_F64, tool, n, s->positions, s->velocities,
                             mass.data(), s->kinematic, targets, double(h), gd, contact.data());
            for (size_t k = 0; k < 3 * size_t(n); ++k)
                s->positions[k] = double(T(s->positions[k]) + T(contact[k]) * (T(0.5) * h));
        }
        std::vector<double> predicted(s->positions, s->positions + 3 * size_t(n));
        fit_rotations<T>(s);
        for (uint32_t pass = 0; pass < cfg->corotation_iterations; ++pass) {
            assemble(sys, s, double(h), predicted.data(), free_mask.data(), applied.data());
            const uint32_t na = uint32_t(sys.active.size());
            std::vector<double> dx(3 * size_t(na));
            uint32_t iters = 0;
            double residual = 0.0;
            if (cfg->exact) {
                std::vector<uint32_t> nbr(6 * size_t(na));
                std::vector<double> blocks(54 * size_t(na)), diag(9 * size_t(na)), rhs(3 * size_t(na));
                export_rows_t(sys, nullptr, nbr.data(), blocks.data(), diag.data(), nullptr, rhs.data());
                residual = vxo_solve_exact(na, rhs.data(), nbr.data(), diag.data(), blocks.data(),
                                           cfg->exact_tolerance, 200000, dx.data(), &iters);
            } else {
                solve_cpu(sys, cfg->precond, cfg->iterations, cfg->solver_relative_tolerance, 1, 0xFFFFFFFFu,
                          dx.data(), &iters, &residual, nullptr);
            }
            for (uint32_t k = 0; k < na; ++k) {
                uint32_t i = sys.active[k];
                for (int a = 0; a < 3; ++a)
                    s->positions[3 * i + a] = double(T(T(s->positions[3 * i + a]) + T(dx[3 * k + a])));
            }
            rep.pcg_iterations += iters;
            rep.passes += 1;
            rep.solver_residual = residual;
            if (pass + 1 == cfg->corotation_iterations) break;
            T angle = fit_rotations<T>(s);
            rep.last_angle = double(angle);
            if (double(angle) <= cfg->corotation_tolerance && residual <= cfg->solver_relative_tolerance) break;
        }
        if (last_force || lambda) bond_forces_t<T>(sys, s, h, last_force, lambda);
        if (fracture) fracture_pass<T>(sys, s, fracture, h);
        // Velocities from the solved positions (no contact): src/solver.rs:668-683.
        for (uint32_t i = 0; i < n; ++i) {
            if (s->kinematic[i]) continue;
            for (int a = 0; a < 3; ++a) {
                T pa = T(s->positions[3 * i + a]);
                if (has_target(i, a)) {
                    s->velocities[3 * i + a] = double((pa - T(previous[3 * i + a])) / h);
                } else {
                    T cv = (pa - T(predicted[3 * i + a])) / h;
                    s->velocities[3 * i + a] = double(drift[i][a] + cv + T(contact[3 * i + a]));
                }
            }
        }
        // The tool's zero-time pass on the final state; its contact velocity
        // is discarded (src/solver.rs:701-706). The tool reports the timed
        // advance's microsteps and both advances' contacts.
        if (tool) {
            const uint32_t steps = tool->microsteps, contacts = tool->contacts;
            std::vector<double> discard(3 * size_t(n), 0.0);
            vxo_tool_advance(std::is_same<T, float>::value ? VXO_F32 : VXO_F64, tool, n, s->positions, s->velocities,
                             mass.data(), s->kinematic, targets, 0.0, gd, discard.data());
            tool->microsteps = steps;
            tool->contacts += contacts;
        }
    }
    if (report) *report = rep;
}

void vxo_step(int mode, vxo_scene *s, const vxo_step_config *cfg, const double *targets, const double *loads,
              vxo_step_report *report, double *last_force, double *lambda) {
    if (mode == VXO_F32)
        step_t<float>(s, cfg, targets, loads, report, last_force, lambda);
    else
        step_t<double>(s, cfg, targets, loads, report, last_force, lambda);
}

void vxo_step_tool(int mode, vxo_scene *s, const vxo_step_config *cfg, const double *targets, const double *loads,
                   vxo_tool *tool, vxo_fracture_state *fracture, vxo_step_report *report) {
    if (mode == VXO_F32)
        step_t<float>(s, cfg, targets, loads, report, nullptr, nullptr, fracture, tool);
    else
        step_t<double>(s, cfg, targets, loads, report, nullptr, nullptr, fracture, tool);
}

void vxo_step_fracture(int mode, vxo_scene *s, const vxo_step_config *cfg, const double *targets, const double *loads,
                       vxo_fracture_state *fracture, vxo_step_report *report, double *last_force, double *lambda) {
    if (mode == VXO_F32)
        step_t<float>(s, cfg, targets, loads, report, last_force, lambda, fracture);
    else
        step_t<double>(s, cfg, targets, loads, report, last_force, lambda, fracture);
}
