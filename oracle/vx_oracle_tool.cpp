// CPU oracle of the rigid tool. TEST INFRASTRUCTURE ONLY: what the tests
// compare the device tool with; never part of the product path.
//
// Restates Tool::advance (src/tool.rs:192-431) with its helpers
// (constrain_pose :156-166, contact_correction :167-175, angular_response
// :181-188, project_translation/linear_response :149-155, key :432-434,
// obb_bin_bounds :438-443, obb_aabb_contact :448-482). Scalar type T: f32 as
// the reference computes, or f64 (the device's precision). Quaternions are
// (x, y, z, w); the vector rotation and the quaternion product follow glam's
// formulas (glam 0.27, Quat::mul_vec3 and Quat::mul_quat).
#include "vx_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <vector>

namespace {

constexpr double kCellT = 0.01;  // src/scene.rs:10

template <class T> struct V {
    T x, y, z;
};
template <class T> V<T> vv(T x, T y, T z) { return V<T>{x, y, z}; }
template <class T> V<T> operator+(V<T> a, V<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V<T> operator-(V<T> a, V<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> V<T> operator*(V<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> V<T> operator/(V<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> V<T> hmul(V<T> a, V<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> T dot(V<T> a, V<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> V<T> cross(V<T> a, V<T> b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
template <class T> T length(V<T> a) { return std::sqrt(dot(a, a)); }
template <class T> V<T> normalize(V<T> a) { return a * (T(1) / length(a)); }
template <class T> V<T> vabs(V<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class T> T element_sum(V<T> a) { return (a.x + a.y) + a.z; }
template <class T> bool is_zero(V<T> a) { return a.x == T(0) && a.y == T(0) && a.z == T(0); }
// glam Vec3::try_normalize: None unless the reciprocal length is finite and positive.
template <class T> V<T> try_normalize_or_zero(V<T> a) {
    const T rcp = T(1) / length(a);
    if (std::isfinite(rcp) && rcp > T(0)) return a * rcp;
    return {T(0), T(0), T(0)};
}
template <class T> T signum(T x) { return std::signbit(x) ? T(-1) : T(1); }

template <class T> struct Q {
    T x, y, z, w;
};
// glam Quat::mul_vec3: v (w^2 - b.b) + b (2 v.b) + (b x v) 2w.
template <class T> V<T> rotate(Q<T> q, V<T> v) {
    const V<T> b{q.x, q.y, q.z};
    const T b2 = dot(b, b);
    return (v * (q.w * q.w - b2) + b * (dot(v, b) * T(2))) + cross(b, v) * (q.w * T(2));
}
template <class T> Q<T> conj(Q<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
// Hamilton product a * b.
template <class T> Q<T> qmul(Q<T> a, Q<T> b) {
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T> Q<T> qnormalize(Q<T> q) {
    const T len = std::sqrt(((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w);
    const T r = T(1) / len;
    return {q.x * r, q.y * r, q.z * r, q.w * r};
}
template <class T> Q<T> from_axis_angle(V<T> axis, T angle) {
    const T s = std::sin(angle * T(0.5)), c = std::cos(angle * T(0.5));
    return {axis.x * s, axis.y * s, axis.z * s, c};
}

template <class T> struct ToolT {
    std::vector<V<T>> cells;
    T mass;
    V<T> inv_inertia, position, velocity, angular, force, torque;
    Q<T> rotation;
    bool mount;
    V<T> m_local, m_world, m_dir;
    Q<T> m_rot;
    T friction;
};

template <class T> V<T> project_translation(const ToolT<T> &t, V<T> v) {
    return t.mount ? t.m_dir * dot(t.m_dir, v) : v;
}
template <class T> V<T> linear_response(const ToolT<T> &t, V<T> impulse) { return project_translation(t, impulse) / t.mass; }
template <class T> V<T> angular_response(const ToolT<T> &t, V<T> torque) {
    if (t.mount) return {T(0), T(0), T(0)};
    return rotate(t.rotation, hmul(t.inv_inertia, rotate(conj(t.rotation), torque)));
}
template <class T> void constrain_pose(ToolT<T> &t) {
    if (!t.mount) return;
    t.rotation = t.m_rot;
    const V<T> base = t.position + rotate(t.rotation, t.m_local);
    t.position = t.position + ((t.m_world + t.m_dir * dot(base - t.m_world, t.m_dir)) - base);
    t.velocity = project_translation(t, t.velocity);
    t.angular = {T(0), T(0), T(0)};
}
template <class T> V<T> contact_correction(const ToolT<T> &t, V<T> normal, T penetration) {
    const V<T> projected = project_translation(t, normal);
    const T mobility = dot(normal, projected);
    if (mobility > T(1e-6)) return projected * (-penetration) / mobility;
    return {T(0), T(0), T(0)};
}

struct Key {
    int32_t k[3];
    bool operator<(const Key &o) const {
        return k[2] != o.k[2] ? k[2] < o.k[2] : (k[1] != o.k[1] ? k[1] < o.k[1] : k[0] < o.k[0]);
    }
};
template <class T> Key key_of(V<T> p) {
    const T c = T(kCellT);
    return {{int32_t(std::floor(p.x / c)), int32_t(std::floor(p.y / c)), int32_t(std::floor(p.z / c))}};
}
template <class T> void bin_bounds(V<T> center, const V<T> (&axes)[3], Key &lo, Key &hi) {
    const T half = T(kCellT) * T(0.5);
    const V<T> ext = (vabs(axes[0]) + vabs(axes[1]) + vabs(axes[2])) * half + vv<T>(half + T(1e-6), half + T(1e-6), half + T(1e-6));
    lo = key_of(center - ext);
    hi = key_of(center + ext);
}
// obb_aabb_contact: the separating-axis test over the tool REV's axes, the
// world axes and their nine cross products; false when any axis separates
// (overlap <= 1e-6), else the axis of least overlap pointing to the voxel.
template <class T> bool obb_aabb(V<T> center, const V<T> (&axes)[3], V<T> voxel, V<T> &normal, T &pen) {
    const T half = T(kCellT) * T(0.5);
    const V<T> world[3] = {vv<T>(1, 0, 0), vv<T>(0, 1, 0), vv<T>(0, 0, 1)};
    const V<T> delta = voxel - center;
    T best = T(INFINITY);
    V<T> bn{T(0), T(0), T(0)};
    V<T> cand[15];
    for (int i = 0; i < 3; ++i) cand[i] = axes[i];
    for (int i = 0; i < 3; ++i) cand[3 + i] = world[i];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cand[6 + 3 * a + b] = cross(axes[a], world[b]);
    for (int c = 0; c < 15; ++c) {
        V<T> axis = cand[c];
        const T len = length(axis);
        if (len < T(1e-6)) continue;
        axis = axis / len;
        const T rt = half * ((std::fabs(dot(axis, axes[0])) + std::fabs(dot(axis, axes[1]))) + std::fabs(dot(axis, axes[2])));
        const T rv = half * element_sum(vabs(axis));
        const T sd = dot(delta, axis);
        const T overlap = rt + rv - std::fabs(sd);
        if (overlap <= T(1e-6)) return false;
        if (overlap < best) {
            best = overlap;
            bn = axis * signum(sd);
        }
    }
    if (is_zero(bn)) return false;
    normal = bn;
    pen = best;
    return true;
}

template <class T>
void advance_t(vxo_tool *io, uint32_t n, const double *positions, double *velocities, const double *mass,
               const uint8_t *kinematic, const double *targets, double dt_in, const double *gravity,
               double *contact_velocity) {
    ToolT<T> t;
    t.cells.resize(io->n_cells);
    for (uint32_t c = 0; c < io->n_cells; ++c)
        t.cells[c] = vv<T>(T(io->cells[3 * c]), T(io->cells[3 * c + 1]), T(io->cells[3 * c + 2]));
    auto v3 = [](const double *p) { return vv<T>(T(p[0]), T(p[1]), T(p[2])); };
    t.mass = T(io->mass);
    t.inv_inertia = v3(io->inverse_inertia);
    t.position = v3(io->position);
    t.rotation = {T(io->rotation[0]), T(io->rotation[1]), T(io->rotation[2]), T(io->rotation[3])};
    t.velocity = v3(io->velocity);
    t.angular = v3(io->angular_velocity);
    t.force = v3(io->external_force);
    t.torque = v3(io->external_torque);
    t.friction = T(io->friction);
    t.mount = io->has_mount != 0;
    t.m_local = v3(io->mount_local_base);
    t.m_world = v3(io->mount_world_base);
    t.m_dir = v3(io->mount_direction);
    t.m_rot = {T(io->mount_rotation[0]), T(io->mount_rotation[1]), T(io->mount_rotation[2]), T(io->mount_rotation[3])};
    const T dt = T(dt_in);
    const T cell = T(kCellT);
    auto has_target = [&](uint32_t i, int a) { return targets && !std::isnan(targets[3 * size_t(i) + a]); };
    auto pos = [&](uint32_t i) { return vv<T>(T(positions[3 * i]), T(positions[3 * i + 1]), T(positions[3 * i + 2])); };
    auto vel = [&](uint32_t i) { return vv<T>(T(velocities[3 * i]), T(velocities[3 * i + 1]), T(velocities[3 * i + 2])); };
    auto set_vel = [&](uint32_t i, V<T> v) {
        velocities[3 * i] = double(v.x);
        velocities[3 * i + 1] = double(v.y);
        velocities[3 * i + 2] = double(v.z);
    };
    std::vector<V<T>> cv(n, vv<T>(0, 0, 0));

    constrain_pose(t);
    T radius = T(0);
    for (const V<T> &p : t.cells) radius = std::max(radius, length(p));
    const V<T> g = v3(gravity);
    const V<T> lin_acc = project_translation(t, g + t.force / t.mass);
    const V<T> ang_acc = angular_response(t, t.torque);
    const T ld = length(t.velocity) * dt + T(0.5) * length(lin_acc) * dt * dt;
    const T ad = (length(t.angular) * dt + T(0.5) * length(ang_acc) * dt * dt) * radius;
    const T ls = ld / (cell * T(0.4)), as = ad / (cell * T(0.4));
    const size_t steps = std::max<size_t>(size_t(std::ceil(std::max(ls, as))), 1);
    const T h = dt / T(steps);
    // Spatial bins of the scene positions, voxel indices in index order; the
    // positions do not change during the advance.
    std::map<Key, std::vector<uint32_t>> spatial;
    for (uint32_t i = 0; i < n; ++i) spatial[key_of(pos(i))].push_back(i);
    uint32_t contacts = 0;
    double friction_heat = 0.0;
    auto free_of = [&](uint32_t i) {
        return vv<T>((!kinematic[i] && !has_target(i, 0)) ? T(1) : T(0), (!kinematic[i] && !has_target(i, 1)) ? T(1) : T(0),
                     (!kinematic[i] && !has_target(i, 2)) ? T(1) : T(0));
    };
    auto for_each_pair = [&](const V<T> (&axes)[3], auto &&fn) {
        for (const V<T> &cell_local : t.cells) {
            const V<T> center = t.position + rotate(t.rotation, cell_local);
            Key lo, hi;
            bin_bounds(center, axes, lo, hi);
            for (int32_t z = lo.k[2]; z <= hi.k[2]; ++z)
                for (int32_t y = lo.k[1]; y <= hi.k[1]; ++y)
                    for (int32_t x = lo.k[0]; x <= hi.k[0]; ++x) {
                        auto it = spatial.find(Key{{x, y, z}});
                        if (it == spatial.end()) continue;
                        for (uint32_t i : it->second) fn(center, i);
                    }
        }
    };
    for (size_t step = 0; step < steps; ++step) {
        t.velocity = t.velocity + lin_acc * h;
        t.angular = t.angular + angular_response(t, t.torque) * h;
        t.position = t.position + t.velocity * h;
        const T angle = length(t.angular) * h;
        if (angle > T(1e-8)) t.rotation = qnormalize(qmul(from_axis_angle(normalize(t.angular), angle), t.rotation));
        const V<T> axes[3] = {rotate(t.rotation, vv<T>(1, 0, 0)), rotate(t.rotation, vv<T>(0, 1, 0)),
                              rotate(t.rotation, vv<T>(0, 0, 1))};
        V<T> correction{T(0), T(0), T(0)};
        T max_pen = T(0);
        for_each_pair(axes, [&](V<T> center, uint32_t i) {
            V<T> nrm;
            T pen;
            if (!obb_aabb(center, axes, pos(i), nrm, pen)) return;
            contacts += 1;
            const V<T> arm = pos(i) - t.position;
            const V<T> rel = t.velocity + cross(t.angular, arm) - vel(i);
            const T closing = dot(rel, nrm);
            const V<T> fr = free_of(i);
            const T mi = T(mass[i]);
            const T wi = !kinematic[i] ? dot(hmul(nrm, fr), nrm) / mi : T(0);
            const T denom = wi + dot(nrm, linear_response(t, nrm)) + dot(nrm, cross(angular_response(t, cross(arm, nrm)), arm));
            if (closing > T(0)) {
                const T normal_impulse = closing / std::max(denom, T(1e-9));
                const V<T> nv = nrm * normal_impulse;
                const V<T> dv = hmul(nv, fr) / mi;
                set_vel(i, vel(i) + dv);
                cv[i] = cv[i] + dv;
                t.velocity = t.velocity - linear_response(t, nv);
                t.angular = t.angular - angular_response(t, cross(arm, nv));
                const V<T> after = t.velocity + cross(t.angular, arm) - vel(i);
                const V<T> tangent = after - nrm * dot(after, nrm);
                const T tspeed = length(tangent);
                const V<T> tdir = try_normalize_or_zero(tangent);
                const T tw = !kinematic[i] ? dot(hmul(tdir, fr), tdir) / mi : T(0);
                const T tden = tw + dot(tdir, linear_response(t, tdir)) + dot(tdir, cross(angular_response(t, cross(arm, tdir)), arm));
                if (tspeed > T(1e-6) && tden > T(1e-9)) {
                    const T fmag = std::min(tspeed / tden, t.friction * normal_impulse);
                    const V<T> f = tdir * fmag;
                    const V<T> fdv = hmul(f, fr) / mi;
                    set_vel(i, vel(i) + fdv);
                    cv[i] = cv[i] + fdv;
                    t.velocity = t.velocity - linear_response(t, f);
                    t.angular = t.angular - angular_response(t, cross(arm, f));
                    const double heat = std::max(double(fmag) * double(tspeed) - 0.5 * double(fmag) * double(fmag) * double(tden), 0.0);
                    friction_heat += heat;
                }
            }
            if (pen > max_pen) {
                max_pen = pen;
                correction = contact_correction(t, nrm, pen);
            }
        });
        t.position = t.position + correction;
        constrain_pose(t);
        for (int pass = 0; pass < 6; ++pass) {
            V<T> corr{T(0), T(0), T(0)};
            T mp = T(1e-6);
            for_each_pair(axes, [&](V<T> center, uint32_t i) {
                V<T> nrm;
                T pen;
                if (!obb_aabb(center, axes, pos(i), nrm, pen)) return;
                if (pen > mp) {
                    mp = pen;
                    corr = contact_correction(t, nrm, pen);
                }
            });
            if (mp <= T(1e-6)) break;
            t.position = t.position + corr;
            constrain_pose(t);
        }
    }
    auto out3 = [](double *d, V<T> v) {
        d[0] = double(v.x);
        d[1] = double(v.y);
        d[2] = double(v.z);
    };
    out3(io->position, t.position);
    io->rotation[0] = double(t.rotation.x);
    io->rotation[1] = double(t.rotation.y);
    io->rotation[2] = double(t.rotation.z);
    io->rotation[3] = double(t.rotation.w);
    out3(io->velocity, t.velocity);
    out3(io->angular_velocity, t.angular);
    io->friction_heat += friction_heat;
    io->microsteps = uint32_t(steps);
    io->contacts = contacts;
    for (uint32_t i = 0; i < n; ++i) out3(contact_velocity + 3 * size_t(i), cv[i]);
}

}  // namespace

void vxo_tool_advance(int mode, vxo_tool *tool, uint32_t n, const double *positions, double *velocities,
                      const double *mass, const uint8_t *kinematic, const double *targets, double dt,
                      const double *gravity, double *contact_velocity) {
    if (mode == VXO_F32)
        advance_t<float>(tool, n, positions, velocities, mass, kinematic, targets, dt, gravity, contact_velocity);
    else
        advance_t<double>(tool, n, positions, velocities, mass, kinematic, targets, dt, gravity, contact_velocity);
}
