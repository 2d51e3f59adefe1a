// Rigid tool on the device: Tool::advance (src/tool.rs:192-431) for the tool
// of a one-scene arena, per substep after the prediction. Included by
// vx_grid.hip (uses its arena planes and tile context).
//
// The reference advances the tool in microsteps of at most 0.4 CELL of
// travel; in each it tests every tool cell against the scene REVs of the
// spatial bins its oriented box touches (15-axis separating-axis test) and
// applies an inelastic normal impulse and Coulomb friction per contact, in
// the order (tool cell, bin z, y, x, REV index), each impulse changing the
// tool velocity that the next contact sees; then it removes the deepest
// overlap and re-projects up to six times. That order is kept here:
//
//  * candidates: the REVs whose predicted positions lie in the tool's body
//    box (cells' bounding box padded by twice the substep's travel and six
//    cells), found by a tile-parallel count / scan / write, in index order;
//  * one workgroup sorts them by bin key (z, y, x, then index), and per
//    microstep finds each cell's contacts in parallel (binary search of the
//    bins, separating-axis test), then thread 0 applies the impulses over the
//    cells and their contacts in the reference's order (the sequential
//    dependency is the tool's velocity), and the projection passes take the
//    first deepest overlap in that order by a workgroup reduction;
//  * the contact velocity of each candidate drifts its position by half a
//    substep (src/solver.rs:535-543) and joins its drift velocity, so the
//    substep's final velocity includes it (src/solver.rs:668-683);
//  * after the substep's velocities, a zero-time advance (no motion, one
//    microstep) applies the impulses and projection once more on the final
//    state and discards its contact velocities (src/solver.rs:701-706).
//
// Arithmetic is FP64 without contraction, the operations in the reference's
// order (glam's vector rotation and quaternion product).

constexpr uint32_t kToolMaxCand = 1u << 18;   // candidate REVs per advance
constexpr uint32_t kToolCellSlots = 32;       // contacts per tool cell per pass
constexpr uint32_t kToolBlock = 512;
constexpr double kToolCell = 0.01;            // src/scene.rs:10
constexpr uint32_t kToolMaxSteps = 1u << 16;  // microsteps per substep (0.4 cells each: 260 m per substep)
enum : uint32_t { TS_CAND = 1, TS_SLOTS = 2, TS_KEY = 4, TS_MARGIN = 8, TS_STEPS = 16 };

struct ToolDev {
    double mass, friction, radius;
    double inv_inertia[3], pos[3], rot[4], vel[3], ang[3], force[3], torque[3];
    double m_local[3], m_world[3], m_dir[3], m_rot[4];
    double pos0[3], rot0[4], box_lo[3], box_hi[3];  // pose at the advance's start, candidate box (body frame)
    double cell_lo[3], cell_hi[3];                   // bounding box of the cells' centres (body frame)
    double h, lin_acc[3], reach;
    double friction_heat;
    // steps: microsteps of the substep's timed advance; micro: of the advance
    // in progress; contacts: summed over the substep's two advances
    uint32_t n_cells, has_mount, steps, micro, candidates, contacts, status;
};
struct ToolContact {
    uint32_t slot, pad;
    double n[3], pen;
};
struct ToolBufs {
    ToolDev *dev;
    const double *cells;    // 3 per cell, body frame about the centre of mass
    uint32_t *tile_count, *tile_off;
    uint64_t *key, *key2;   // candidates, then sorted
    uint32_t *idx, *idx2;
    double *cv;             // contact velocity, 3 per sorted candidate
    ToolContact *con;       // kToolCellSlots per cell
    uint32_t *ncon;         // per cell
    double *vsave;          // velocities at the advance's start, 3 per sorted candidate
};

struct t3 {
    double x, y, z;
};
__device__ __forceinline__ t3 tv(double x, double y, double z) { return t3{x, y, z}; }
__device__ __forceinline__ t3 tld(const double *p) { return t3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void tst(double *p, t3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
__device__ __forceinline__ t3 operator+(t3 a, t3 b) {
#pragma clang fp contract(off)
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ t3 operator-(t3 a, t3 b) {
#pragma clang fp contract(off)
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ t3 operator*(t3 a, double s) {
#pragma clang fp contract(off)
    return {a.x * s, a.y * s, a.z * s};
}
__device__ __forceinline__ t3 tdiv(t3 a, double s) {
#pragma clang fp contract(off)
    return {a.x / s, a.y / s, a.z / s};
}
__device__ __forceinline__ t3 thmul(t3 a, t3 b) {
#pragma clang fp contract(off)
    return {a.x * b.x, a.y * b.y, a.z * b.z};
}
__device__ __forceinline__ double tdot(t3 a, t3 b) {
#pragma clang fp contract(off)
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
__device__ __forceinline__ t3 tcross(t3 a, t3 b) {
#pragma clang fp contract(off)
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ double tlen(t3 a) {
#pragma clang fp contract(off)
    return sqrt(tdot(a, a));
}
__device__ __forceinline__ t3 tabs(t3 a) { return {fabs(a.x), fabs(a.y), fabs(a.z)}; }
struct tq {
    double x, y, z, w;
};
__device__ __forceinline__ tq tqld(const double *p) { return tq{p[0], p[1], p[2], p[3]}; }
__device__ __forceinline__ void tqst(double *p, tq q) {
    p[0] = q.x;
    p[1] = q.y;
    p[2] = q.z;
    p[3] = q.w;
}
// glam Quat::mul_vec3
__device__ __forceinline__ t3 trot(tq q, t3 v) {
#pragma clang fp contract(off)
    const t3 b{q.x, q.y, q.z};
    const double b2 = tdot(b, b);
    return (v * (q.w * q.w - b2) + b * (tdot(v, b) * 2.0)) + tcross(b, v) * (q.w * 2.0);
}
__device__ __forceinline__ tq tconj(tq q) { return {-q.x, -q.y, -q.z, q.w}; }

__device__ void tool_constrain(ToolDev &T) {
#pragma clang fp contract(off)
    if (!T.has_mount) return;
    for (int k = 0; k < 4; ++k) T.rot[k] = T.m_rot[k];
    const t3 dir = tld(T.m_dir), world = tld(T.m_world);
    const t3 base = tld(T.pos) + trot(tqld(T.rot), tld(T.m_local));
    tst(T.pos, tld(T.pos) + ((world + dir * tdot(base - world, dir)) - base));
    tst(T.vel, dir * tdot(dir, tld(T.vel)));
    tst(T.ang, tv(0, 0, 0));
}
__device__ __forceinline__ t3 tool_project(const ToolDev &T, t3 v) {
#pragma clang fp contract(off)
    if (!T.has_mount) return v;
    const t3 d = tld(T.m_dir);
    return d * tdot(d, v);
}
__device__ __forceinline__ t3 tool_linear(const ToolDev &T, t3 impulse) { return tdiv(tool_project(T, impulse), T.mass); }
__device__ __forceinline__ t3 tool_angular(const ToolDev &T, t3 torque) {
#pragma clang fp contract(off)
    if (T.has_mount) return tv(0, 0, 0);
    const tq q = tqld(T.rot);
    return trot(q, thmul(tld(T.inv_inertia), trot(tconj(q), torque)));
}
__device__ __forceinline__ t3 tool_correction(const ToolDev &T, t3 normal, double pen) {
#pragma clang fp contract(off)
    const t3 projected = tool_project(T, normal);
    const double mobility = tdot(normal, projected);
    if (mobility > 1e-6) return tdiv(projected * (-pen), mobility);
    return tv(0, 0, 0);
}

// Spatial bin key floor(p / CELL) per axis, packed (z, y, x) into 21 bits
// each (biased by 2^20); false outside that range.
__device__ __forceinline__ bool tool_key(t3 p, int32_t (&k)[3]) {
    const double c[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
        const double f = floor(c[a] / kToolCell);
        if (!(f >= -1048576.0 && f < 1048576.0)) return false;
        k[a] = int32_t(f);
    }
    return true;
}
__device__ __forceinline__ uint64_t tool_pack(int32_t x, int32_t y, int32_t z) {
    return (uint64_t(uint32_t(z + 1048576)) << 42) | (uint64_t(uint32_t(y + 1048576)) << 21) | uint64_t(uint32_t(x + 1048576));
}

// obb_aabb_contact (src/tool.rs:448-482).
__device__ bool tool_sat(t3 center, const t3 (&axes)[3], t3 voxel, t3 &normal, double &pen) {
#pragma clang fp contract(off)
    const double half = kToolCell * 0.5;
    const t3 delta = voxel - center;
    double best = INFINITY;
    t3 bn{0, 0, 0};
    for (int c = 0; c < 15; ++c) {
        t3 axis;
        if (c < 3)
            axis = axes[c];
        else if (c < 6)
            axis = tv(c == 3 ? 1.0 : 0.0, c == 4 ? 1.0 : 0.0, c == 5 ? 1.0 : 0.0);
        else {
            const int a = (c - 6) / 3, b = (c - 6) % 3;
            axis = tcross(axes[a], tv(b == 0 ? 1.0 : 0.0, b == 1 ? 1.0 : 0.0, b == 2 ? 1.0 : 0.0));
        }
        const double len = tlen(axis);
        if (len < 1e-6) continue;
        axis = tdiv(axis, len);
        const double rt = half * ((fabs(tdot(axis, axes[0])) + fabs(tdot(axis, axes[1]))) + fabs(tdot(axis, axes[2])));
        const t3 aa = tabs(axis);
        const double rv = half * ((aa.x + aa.y) + aa.z);
        const double sd = tdot(delta, axis);
        const double overlap = rt + rv - fabs(sd);
        if (overlap <= 1e-6) return false;
        if (overlap < best) {
            best = overlap;
            bn = axis * (signbit(sd) ? -1.0 : 1.0);
        }
    }
    if (bn.x == 0.0 && bn.y == 0.0 && bn.z == 0.0) return false;
    normal = bn;
    pen = best;
    return true;
}

// obb_bin_bounds (src/tool.rs:438-443).
__device__ __forceinline__ bool tool_bounds(t3 center, const t3 (&axes)[3], int32_t (&lo)[3], int32_t (&hi)[3]) {
#pragma clang fp contract(off)
    const double half = kToolCell * 0.5;
    const t3 e = (tabs(axes[0]) + tabs(axes[1])) + tabs(axes[2]);
    const double pad = half + 1e-6;
    const t3 ext = e * half + tv(pad, pad, pad);
    return tool_key(center - ext, lo) && tool_key(center + ext, hi);
}

// Start of a step (one thread): the status bits stay set through the step's
// advances (an advance with a bit set does nothing), so that a failure of any
// substep's advance is the step's; vx_scene_set_tool starts from zero too.
__global__ void tool_reset(ToolBufs B) { B.dev->status = 0; }

// Start of an advance (one thread): constrain the pose, the microsteps of
// the substep (src/tool.rs:198-214), the candidate box. `timed` is 0 for the
// zero-time pass (dt = 0).
__global__ void tool_prepare(ToolBufs B, double dt, double gx, double gy, double gz, uint32_t timed) {
#pragma clang fp contract(off)
    ToolDev &T = *B.dev;
    tool_constrain(T);
    const t3 lin = tool_project(T, tv(gx, gy, gz) + tdiv(tld(T.force), T.mass));
    const t3 angacc = tool_angular(T, tld(T.torque));
    const double ld = tlen(tld(T.vel)) * dt + 0.5 * tlen(lin) * dt * dt;
    const double ad = (tlen(tld(T.ang)) * dt + 0.5 * tlen(angacc) * dt * dt) * T.radius;
    const double ls = ld / (kToolCell * 0.4), as = ad / (kToolCell * 0.4);
    const double s = ceil(fmax(ls, as));
    uint32_t steps = s >= 1.0 ? (s < 4294967295.0 ? uint32_t(s) : 4294967295u) : 1u;
    if (steps > kToolMaxSteps) {  // reported after the step; the advances do nothing until then
        T.status |= TS_STEPS;
        steps = 1;
    }
    T.micro = steps;
    if (timed) {
        T.steps = steps;
        T.contacts = 0;
    }
    T.h = dt / double(steps);
    tst(T.lin_acc, lin);
    for (int k = 0; k < 3; ++k) T.pos0[k] = T.pos[k];
    for (int k = 0; k < 4; ++k) T.rot0[k] = T.rot[k];
    // A scene REV that some bin of some cell holds during the advance lies
    // within 4.2 cells of that cell's centre, and no cell moves further than
    // the travel ld + ad plus the overlap corrections (bounded after the
    // advance by `reach`).
    T.reach = 2.0 * (ld + ad) + 2.0 * kToolCell;
    const double pad = T.reach + 4.2 * kToolCell;
    for (int k = 0; k < 3; ++k) {
        T.box_lo[k] = T.cell_lo[k] - pad;
        T.box_hi[k] = T.cell_hi[k] + pad;
    }
    T.candidates = 0;
}

__device__ __forceinline__ bool tool_in_box(const ToolDev &T, t3 p) {
#pragma clang fp contract(off)
    const t3 b = trot(tconj(tqld(T.rot0)), p - tld(T.pos0));
    return b.x >= T.box_lo[0] && b.x <= T.box_hi[0] && b.y >= T.box_lo[1] && b.y <= T.box_hi[1] && b.z >= T.box_lo[2] &&
           b.z <= T.box_hi[2];
}
