// CPU oracle of the voxel contact projection (TEST INFRASTRUCTURE ONLY: the
// HIP path is compared against it; the product library never links or calls
// it).
//
// Restates collision::project (src/collision.rs:33-123) in f32 with the
// reference's order of operations, no fused multiply-adds (the Makefile
// builds with -ffp-contract=off, as Rust never contracts):
//   adjacency            src/collision.rs:506-518
//   candidate_pairs      src/collision.rs:376-430 (hashed bins, a set of
//                        unique pairs, then sorted), bin_bounds :432-439,
//                        ensure_static_bins :462-504 (rebuilt each call)
//   local_continuum_pair src/collision.rs:520-550
//   swept_normal         src/collision.rs:552-596
//   overlap_normal       src/collision.rs:598-621
//   touching_normal      src/collision.rs:441-460
//   normal_axis          src/collision.rs:623-631
//   inverse_mass         src/collision.rs:633-640
//   separate             src/collision.rs:642-665
// The contacts are sorted by (i, j) and one contact is kept per pair: the
// first one found. The reference sorts with sort_unstable_by_key, which
// leaves the choice among a pair's contacts from different passes
// unspecified; they differ only if the pair's normal changed between passes.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "vx_oracle.h"

namespace {

constexpr float kCell = 0.01f;  // src/scene.rs:10
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

struct V3 {
    float v[3];
    float &operator[](int a) { return v[a]; }
    float operator[](int a) const { return v[a]; }
};
V3 load(const float *p, size_t i) { return V3{{p[3 * i], p[3 * i + 1], p[3 * i + 2]}}; }
void store(float *p, size_t i, const V3 &x) {
    p[3 * i] = x[0];
    p[3 * i + 1] = x[1];
    p[3 * i + 2] = x[2];
}
V3 sub(const V3 &a, const V3 &b) { return V3{{a[0] - b[0], a[1] - b[1], a[2] - b[2]}}; }
V3 vabs(const V3 &a) { return V3{{std::fabs(a[0]), std::fabs(a[1]), std::fabs(a[2])}}; }
// glam Vec3::dot: (x*x' + y*y') + z*z'
float dot(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// f32::min / f32::max: the non-NaN operand wins, as fmin / fmax.
float fmin32(float a, float b) { return std::fmin(a, b); }
float fmax32(float a, float b) { return std::fmax(a, b); }

struct Scene {
    uint32_t n;
    float *pos, *prev;
    const float *rest, *mass;
    const uint8_t *kin, *targets;
    std::vector<uint32_t> adj;  // [voxel][6]: -x, +x, -y, +y, -z, +z
};

// inverse_mass (src/collision.rs:633-640)
float inverse_mass(const Scene &s, uint32_t i, int axis) {
    if (s.kin[i] || (s.targets && ((s.targets[i] >> axis) & 1))) return 0.0f;
    return 1.0f / s.mass[i];
}

// bin_bounds (src/collision.rs:432-439): the margin 0.5 + 2e-4 is one f32.
void bin_bounds(const V3 &a, const V3 &b, int lo[3], int hi[3]) {
    const float m = 0.5f + 2e-4f;
    for (int c = 0; c < 3; ++c) {
        lo[c] = int(std::floor(a[c] - m));
        hi[c] = int(std::floor(b[c] + m));
    }
}

void swept_box(const Scene &s, const float *starts, uint32_t i, int lo[3], int hi[3]) {
    const V3 st = load(starts, i), p = load(s.pos, i);
    V3 a, b;
    for (int c = 0; c < 3; ++c) {
        a[c] = fmin32(st[c], p[c]) / kCell;
        b[c] = fmax32(st[c], p[c]) / kCell;
    }
    bin_bounds(a, b, lo, hi);
}

struct KeyHash {
    size_t operator()(const std::array<int, 3> &k) const {
        return (size_t(uint32_t(k[0])) * 73856093u) ^ (size_t(uint32_t(k[1])) * 19349663u) ^
               (size_t(uint32_t(k[2])) * 83492791u);
    }
};
using Bins = std::unordered_map<std::array<int, 3>, std::vector<uint32_t>, KeyHash>;

void add_to_bins(Bins &bins, uint32_t i, const int lo[3], const int hi[3]) {
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) bins[{x, y, z}].push_back(i);
}

// candidate_pairs (src/collision.rs:376-430) with ensure_static_bins built
// fresh: dynamic voxels into `bins`, kinematic voxels into `static_bins`,
// every unordered pair sharing a dynamic bin plus every (dynamic, kinematic)
// pair sharing a bin, each once, sorted.
std::vector<std::pair<uint32_t, uint32_t>> candidate_pairs(const Scene &s, const float *starts) {
    Bins bins, static_bins;
    int lo[3], hi[3];
    for (uint32_t i = 0; i < s.n; ++i) {
        swept_box(s, starts, i, lo, hi);
        add_to_bins(s.kin[i] ? static_bins : bins, i, lo, hi);
    }
    std::unordered_set<uint64_t> unique;
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    auto add = [&](uint32_t a, uint32_t b) {
        const uint32_t i = std::min(a, b), j = std::max(a, b);
        if (unique.insert((uint64_t(i) << 32) | j).second) pairs.emplace_back(i, j);
    };
    const char *env = std::getenv("VXO_PROGRESS");
    const bool progress = env && env[0] == '1';
    size_t done = 0;
    for (const auto &kv : bins) {
        const std::vector<uint32_t> &ids = kv.second;
        for (size_t a = 0; a < ids.size(); ++a)
            for (size_t b = a + 1; b < ids.size(); ++b) add(ids[a], ids[b]);
        auto f = static_bins.find(kv.first);
        if (f != static_bins.end())
            for (uint32_t i : ids)
                for (uint32_t j : f->second) add(i, j);
        if (progress && ++done % (bins.size() / 16 + 1) == 0) {
            std::fprintf(stderr, "oracle contact: %zu of %zu bins paired, %zu pairs\n", done, bins.size(), pairs.size());
            std::fflush(stderr);
        }
    }
    std::sort(pairs.begin(), pairs.end());
    return pairs;
}

// adjacency (src/collision.rs:506-518)
void build_adjacency(Scene &s, const uint32_t *bond_neighbor) {
    s.adj.assign(size_t(s.n) * 6, kEmpty);
    for (int axis = 0; axis < 3; ++axis)
        for (uint32_t i = 0; i < s.n; ++i) {
            const uint32_t j = bond_neighbor[size_t(axis) * s.n + i];
            if (j == kEmpty) continue;
            s.adj[size_t(i) * 6 + 2 * axis + 1] = j;
            s.adj[size_t(j) * 6 + 2 * axis] = i;
        }
}

// connected (src/collision.rs:529-549): a monotone path of intact bonds
// inside the lattice cell, one axis step at a time.
bool connected(const Scene &s, uint32_t i, uint32_t j, V3 steps) {
    if (steps[0] == 0.0f && steps[1] == 0.0f && steps[2] == 0.0f) return i == j;
    for (int axis = 0; axis < 3; ++axis) {
        if (steps[axis] == 0.0f) continue;
        const uint32_t next = s.adj[size_t(i) * 6 + 2 * axis + (steps[axis] > 0.0f ? 1 : 0)];
        if (next == kEmpty) continue;
        V3 remaining = steps;
        remaining[axis] = 0.0f;
        if (connected(s, next, j, remaining)) return true;
    }
    return false;
}

// local_continuum_pair (src/collision.rs:520-550)
bool local_continuum_pair(const Scene &s, uint32_t i, uint32_t j) {
    const V3 ri = load(s.rest, i), rj = load(s.rest, j);
    V3 delta, steps;
    float frac = 0.0f, big = 0.0f;
    for (int c = 0; c < 3; ++c) {
        delta[c] = (rj[c] - ri[c]) / kCell;
        steps[c] = std::round(delta[c]);  // half away from zero, as f32::round
    }
    // max_element is x.max(y.max(z))
    frac = fmax32(std::fabs(delta[0] - steps[0]), fmax32(std::fabs(delta[1] - steps[1]), std::fabs(delta[2] - steps[2])));
    big = fmax32(std::fabs(steps[0]), fmax32(std::fabs(steps[1]), std::fabs(steps[2])));
    if (frac > 1e-3f || big > 1.0f) return false;
    return connected(s, i, j, steps);
}

// swept_normal (src/collision.rs:552-596); returns the axis, or -1.
int swept_normal(const V3 &si, const V3 &sj, const V3 &ei, const V3 &ej, float *sign_out) {
    const V3 d0 = sub(sj, si), di = sub(ei, si), dj = sub(ej, sj), rel = sub(dj, di);
    const V3 a0 = vabs(d0);
    const float close = kCell - 1e-6f;
    if (a0[0] <= close && a0[1] <= close && a0[2] <= close) return -1;
    float enter = 0.0f, exit = 1.0f, sign = 1.0f;
    int hit = -1;
    for (int axis = 0; axis < 3; ++axis) {
        if (std::fabs(rel[axis]) < 1e-12f) {
            if (std::fabs(d0[axis]) > kCell) return -1;
            continue;
        }
        const float t0 = (-kCell - d0[axis]) / rel[axis];
        const float t1 = (kCell - d0[axis]) / rel[axis];
        const float near_t = fmin32(t0, t1), far_t = fmax32(t0, t1);
        if (near_t > enter || (hit < 0 && near_t >= enter)) {
            enter = near_t;
            hit = axis;
            sign = (d0[axis] + rel[axis] * enter >= 0.0f) ? 1.0f : -1.0f;
        }
        exit = fmin32(exit, far_t);
        if (enter > exit) return -1;
    }
    if (exit < 0.0f || enter > 1.0f) return -1;
    *sign_out = sign;
    return hit;
}

// overlap_normal (src/collision.rs:598-621); returns the axis, or -1.
int overlap_normal(const V3 &a, const V3 &b, const V3 &fallback, float *sign_out, float *depth) {
    const V3 delta = sub(b, a), ad = vabs(delta);
    const V3 ov{{kCell - ad[0], kCell - ad[1], kCell - ad[2]}};
    if (fmin32(ov[0], fmin32(ov[1], ov[2])) <= 0.0f) return -1;
    const int axis = (ov[0] <= ov[1] && ov[0] <= ov[2]) ? 0 : (ov[1] <= ov[2] ? 1 : 2);
    float src = 1.0f;
    if (std::fabs(delta[axis]) > 1e-8f)
        src = delta[axis];
    else if (std::fabs(fallback[axis]) > 1e-8f)
        src = fallback[axis];
    *sign_out = std::signbit(src) ? -1.0f : 1.0f;  // f32::signum of a nonzero value
    *depth = ov[axis];
    return axis;
}

// touching_normal (src/collision.rs:441-460); returns the axis, or -1.
int touching_normal(const V3 &a, const V3 &b, float *sign_out) {
    const float slop = 2.0e-6f;
    const V3 delta = sub(b, a), ad = vabs(delta);
    const V3 gaps{{ad[0] - kCell, ad[1] - kCell, ad[2] - kCell}};
    const int axis = (gaps[0] >= gaps[1] && gaps[0] >= gaps[2]) ? 0 : (gaps[1] >= gaps[2] ? 1 : 2);
    if (gaps[axis] > slop) return -1;
    for (int o = 0; o < 3; ++o)
        if (o != axis && gaps[o] >= -slop) return -1;
    *sign_out = delta[axis] >= 0.0f ? 1.0f : -1.0f;
    return axis;
}

V3 axis_normal(int axis, float sign) {
    V3 n{{0.0f, 0.0f, 0.0f}};
    n[axis] = sign;
    return n;
}

// separate (src/collision.rs:642-665); normal_axis of an axis normal is its axis.
bool separate(Scene &s, uint32_t i, uint32_t j, const V3 &normal, int axis, float amount) {
    const float wi = inverse_mass(s, i, axis), wj = inverse_mass(s, j, axis);
    const float denom = wi + wj;
    if (denom <= 0.0f) return false;
    const float si = amount * wi / denom, sj = amount * wj / denom;
    V3 pi = load(s.pos, i), pj = load(s.pos, j), qi = load(s.prev, i), qj = load(s.prev, j);
    for (int c = 0; c < 3; ++c) {
        const float di = -normal[c] * si, dj = normal[c] * sj;
        pi[c] += di;
        pj[c] += dj;
        qi[c] += di;
        qj[c] += dj;
    }
    store(s.pos, i, pi);
    store(s.pos, j, pj);
    store(s.prev, i, qi);
    store(s.prev, j, qj);
    return true;
}

struct Contact {
    uint32_t i, j;
    int axis;
    float sign;
    bool new_impact;
};

}  // namespace

extern "C" int vxo_local_continuum_pair(uint32_t n, const float *rest, const uint32_t *bond_neighbor, uint32_t i,
                                        uint32_t j) {
    Scene s{};
    s.n = n;
    s.rest = rest;
    build_adjacency(s, bond_neighbor);
    return local_continuum_pair(s, i, j) ? 1 : 0;
}

extern "C" uint64_t vxo_contact_candidates(uint32_t n, const float *positions, const float *previous,
                                           const uint8_t *kinematic, uint32_t *out_pairs, uint64_t capacity) {
    Scene s{};
    s.n = n;
    s.pos = const_cast<float *>(positions);
    s.kin = kinematic;
    const auto pairs = candidate_pairs(s, previous);
    for (size_t k = 0; k < pairs.size() && k < capacity; ++k) {
        out_pairs[2 * k] = pairs[k].first;
        out_pairs[2 * k + 1] = pairs[k].second;
    }
    return pairs.size();
}

extern "C" uint64_t vxo_contact_project(uint32_t n, float *positions, float *previous, const float *rest,
                                        const float *mass, const uint8_t *kinematic, const uint8_t *targets,
                                        const uint32_t *bond_neighbor, uint64_t capacity, uint32_t *out_i,
                                        uint32_t *out_j, float *out_normal, uint8_t *out_new_impact,
                                        uint64_t *stats) {
    Scene s{};
    s.n = n;
    s.pos = positions;
    s.prev = previous;
    s.rest = rest;
    s.mass = mass;
    s.kin = kinematic;
    s.targets = targets;
    // let starts = scene.previous.clone()
    const std::vector<float> starts(previous, previous + size_t(n) * 3);
    build_adjacency(s, bond_neighbor);
    auto pairs = candidate_pairs(s, starts.data());
    const uint64_t candidates = pairs.size();
    pairs.erase(std::remove_if(pairs.begin(), pairs.end(),
                               [&](const std::pair<uint32_t, uint32_t> &p) {
                                   return local_continuum_pair(s, p.first, p.second);
                               }),
                pairs.end());
    // VXO_PROGRESS=1: a line on stderr after the broad phase and each pass
    // (long single-thread runs on large states).
    const char *env = std::getenv("VXO_PROGRESS");
    const bool progress = env && env[0] == '1';
    if (progress) {
        std::fprintf(stderr, "oracle contact: %llu candidate pairs, %zu after the local-continuum filter\n",
                     (unsigned long long)candidates, pairs.size());
        std::fflush(stderr);
    }
    std::vector<Contact> contacts;
    uint64_t passes = 0;
    for (int pass = 0; pass < 6; ++pass) {
        ++passes;
        bool changed = false;
        for (const auto &p : pairs) {
            const uint32_t i = p.first, j = p.second;
            const V3 si = load(starts.data(), i), sj = load(starts.data(), j);
            const V3 start_delta = sub(sj, si);
            float sign = 1.0f, depth = 0.0f;
            int axis = swept_normal(si, sj, load(s.pos, i), load(s.pos, j), &sign);
            if (axis >= 0) {
                const V3 normal = axis_normal(axis, sign);
                const float separation = dot(sub(load(s.pos, j), load(s.pos, i)), normal);
                if (separation < kCell) {
                    const float correction = kCell - separation + 1e-6f;
                    changed |= separate(s, i, j, normal, axis, correction);
                }
                contacts.push_back({i, j, axis, sign, dot(start_delta, normal) > kCell + 2e-6f});
                continue;
            }
            axis = overlap_normal(load(s.pos, i), load(s.pos, j), start_delta, &sign, &depth);
            if (axis >= 0) {
                const V3 normal = axis_normal(axis, sign);
                changed |= separate(s, i, j, normal, axis, depth + 1e-6f);
                contacts.push_back({i, j, axis, sign, dot(start_delta, normal) > kCell + 2e-6f});
                continue;
            }
            axis = touching_normal(load(s.pos, i), load(s.pos, j), &sign);
            if (axis >= 0) contacts.push_back({i, j, axis, sign, false});
        }
        if (progress) {
            std::fprintf(stderr, "oracle contact: pass %llu done, %zu contacts so far\n", (unsigned long long)passes,
                         contacts.size());
            std::fflush(stderr);
        }
        if (!changed) break;
    }
    std::stable_sort(contacts.begin(), contacts.end(), [](const Contact &a, const Contact &b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    contacts.erase(std::unique(contacts.begin(), contacts.end(),
                               [](const Contact &a, const Contact &b) { return a.i == b.i && a.j == b.j; }),
                   contacts.end());
    for (size_t k = 0; k < contacts.size() && k < capacity; ++k) {
        out_i[k] = contacts[k].i;
        out_j[k] = contacts[k].j;
        for (int c = 0; c < 3; ++c) out_normal[3 * k + c] = axis_normal(contacts[k].axis, contacts[k].sign)[c];
        out_new_impact[k] = contacts[k].new_impact ? 1 : 0;
    }
    if (stats) {
        stats[0] = candidates;
        stats[1] = pairs.size();
        stats[2] = passes;
        stats[3] = contacts.size();
    }
    return contacts.size();
}
