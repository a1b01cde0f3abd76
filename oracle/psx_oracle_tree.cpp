// psx_oracle_tree.cpp - light tree construction for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY: a C++ restatement (it needs std::nth_element,
// whose element arrangement fixes the emitter order inside leaves) of
// build_cycles_light_tree (src/sampling/light_tree.cpp:90-408,
// include/psycles/sampling/light_tree.h:15-87): 12-bucket surface-area-
// orientation splits of the local emitters, distant emitters in a dedicated
// branch at the root.  Used by psx_oracle.c through the C interface in
// psx_oracle.h; the product never links it.
#include "psx_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kHalfPi = 0.5f * kPi;
constexpr float kTwoPi = 2.0f * kPi;
constexpr int kBuckets = 12;
constexpr uint32_t kInvalid = 0xffffffffu;

struct Vec {
    float x, y, z;
};
Vec vadd(Vec a, Vec b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec vsub(Vec a, Vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec vmul(Vec a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float vdot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float vlen(Vec a) { return std::sqrt(std::max(vdot(a, a), 0.0f)); }
Vec vnorm_or(Vec v, Vec fallback) {
    const float m = vlen(v);
    return m > 1.0e-20f ? vmul(v, 1.0f / m) : fallback;
}
float comp(Vec v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }
bool finite3(Vec v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

struct Bounds {
    Vec lo{}, hi{};
    bool empty = true;
};
struct Cone {
    Vec axis{};
    float theta_o = 0.0f, theta_e = 0.0f;
    bool empty = true;
};
struct Measure {
    Bounds bounds;
    Cone cone;
    float energy = 0.0f;
};
struct Emitter {
    Measure m;
    Vec centroid{};
    uint32_t id = 0;
    bool distant = false;
};
struct Node {
    Measure m;
    uint32_t parent = kInvalid, left = kInvalid, right = kInvalid;
    uint32_t first = 0, count = 0;
    uint32_t kind = ORACLE_TREE_LEAF;
};

Bounds merge(const Bounds &a, const Bounds &b) {
    if (a.empty) return b;
    if (b.empty) return a;
    Bounds r;
    r.lo = {std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z)};
    r.hi = {std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z)};
    r.empty = false;
    return r;
}

// orientation cone union (light_tree.cpp:296-352)
Cone merge(const Cone &a, const Cone &b) {
    if (a.empty) return b;
    if (b.empty) return a;
    const Cone *wide = &a, *narrow = &b;
    if (b.theta_o > a.theta_o) std::swap(wide, narrow);
    const float c = std::clamp(vdot(wide->axis, narrow->axis), -1.0f, 1.0f);
    const float axis_angle = std::acos(c);
    const float theta_e = std::max(a.theta_e, b.theta_e);
    if (wide->theta_o + 5.0e-4f >= std::min(kPi, axis_angle + narrow->theta_o))
        return Cone{wide->axis, wide->theta_o, theta_e, false};
    const float theta_o = 0.5f * (axis_angle + wide->theta_o + narrow->theta_o);
    if (theta_o >= kPi) return Cone{wide->axis, kPi, theta_e, false};
    const float rotation = theta_o - wide->theta_o;
    Vec ortho = vsub(narrow->axis, vmul(wide->axis, c));
    if (c < -0.9995f) {
        const Vec w = wide->axis;
        const Vec cand = std::abs(w.x) < 0.9f ? Vec{1.0f, 0.0f, 0.0f} : Vec{0.0f, 1.0f, 0.0f};
        ortho = vnorm_or({w.y * cand.z - w.z * cand.y, w.z * cand.x - w.x * cand.z, w.x * cand.y - w.y * cand.x},
                         {0.0f, 0.0f, 1.0f});
    } else {
        ortho = vnorm_or(ortho, {0.0f, 0.0f, 1.0f});
    }
    const Vec axis = vnorm_or(vadd(vmul(wide->axis, std::cos(rotation)), vmul(ortho, std::sin(rotation))), wide->axis);
    return Cone{axis, theta_o, theta_e, false};
}

Measure merge(const Measure &a, const Measure &b) {
    if (!(a.energy > 0.0f)) return b;
    if (!(b.energy > 0.0f)) return a;
    return Measure{merge(a.bounds, b.bounds), merge(a.cone, b.cone), a.energy + b.energy};
}

// surface-area-orientation cost (light_tree.cpp:369-398)
float cost(const Measure &m) {
    if (!(m.energy > 0.0f) || m.cone.empty) return 0.0f;
    float spatial = 0.0f;
    if (!m.bounds.empty) {
        const Vec e = vsub(m.bounds.hi, m.bounds.lo);
        spatial = 2.0f * (e.x * e.y + e.x * e.z + e.y * e.z);
        if (spatial == 0.0f) spatial = vlen(e);
    }
    const float theta_w = std::min(kPi, m.cone.theta_o + m.cone.theta_e);
    const float co = std::cos(m.cone.theta_o), so = std::sin(m.cone.theta_o);
    const float orientation = kTwoPi * (1.0f - co) + kHalfPi * (2.0f * theta_w * so -
                                                                 std::cos(m.cone.theta_o - 2.0f * theta_w) -
                                                                 2.0f * m.cone.theta_o * so + co);
    return m.energy * spatial * orientation;
}

Measure sanitized(const Measure &in) {
    if (!std::isfinite(in.energy) || in.energy <= 0.0f) return Measure{};
    Measure r = in;
    if (!r.bounds.empty && (!finite3(r.bounds.lo) || !finite3(r.bounds.hi))) r.bounds = Bounds{};
    if (!r.cone.empty) {
        r.cone.axis = vnorm_or(r.cone.axis, {0.0f, 0.0f, 1.0f});
        r.cone.theta_o = std::clamp(r.cone.theta_o, 0.0f, kPi);
        r.cone.theta_e = std::clamp(r.cone.theta_e, 0.0f, kPi);
    }
    return r;
}

struct Bucket {
    Measure m;
    uint32_t count = 0;
};

struct Builder {
    std::vector<Emitter> emitters;
    std::vector<Node> nodes;
    std::vector<uint32_t> to_tree, to_leaf;
    uint32_t max_leaf = 8;

    uint32_t add_node(uint32_t parent) {
        Node n;
        n.parent = parent;
        nodes.push_back(n);
        return static_cast<uint32_t>(nodes.size() - 1u);
    }

    uint32_t build(uint32_t begin, uint32_t end, uint32_t parent) {
        const uint32_t index = add_node(parent);
        const uint32_t count = end - begin;
        if (count == 0u) return index;
        Measure measure;
        Bounds centroids;
        for (uint32_t i = begin; i < end; ++i) {
            measure = merge(measure, emitters[i].m);
            Bounds p;
            p.lo = p.hi = emitters[i].centroid;
            p.empty = false;
            centroids = merge(centroids, p);
        }
        nodes[index].m = measure;
        int axis = -1;
        uint32_t middle = begin + count / 2u;
        float best = std::numeric_limits<float>::infinity();
        const float parent_cost = cost(measure);
        const Vec extent = vsub(centroids.hi, centroids.lo);
        const float max_extent = std::max({extent.x, extent.y, extent.z});
        if (count > 1u && parent_cost > 0.0f && max_extent > 0.0f) {
            for (int a = 0; a < 3; ++a) {
                const float ext = comp(extent, a);
                if (!(ext > 0.0f)) continue;
                std::array<Bucket, kBuckets> buckets{};
                for (uint32_t i = begin; i < end; ++i) {
                    const float rel = (comp(emitters[i].centroid, a) - comp(centroids.lo, a)) / ext;
                    const size_t b = std::min<size_t>(
                        static_cast<size_t>(std::max(rel, 0.0f) * static_cast<float>(kBuckets)), kBuckets - 1u);
                    buckets[b].m = merge(buckets[b].m, emitters[i].m);
                    ++buckets[b].count;
                }
                // prefix unions from the left, suffix unions from the right
                std::array<Bucket, kBuckets - 1> left{}, right{};
                left[0] = buckets[0];
                for (int b = 1; b + 1 < kBuckets; ++b)
                    left[b] = Bucket{merge(left[b - 1].m, buckets[b].m), left[b - 1].count + buckets[b].count};
                right[kBuckets - 2] = buckets[kBuckets - 1];
                for (int b = kBuckets - 3; b >= 0; --b)
                    right[b] = Bucket{merge(right[b + 1].m, buckets[b + 1].m), right[b + 1].count + buckets[b + 1].count};
                const float regularization = max_extent / ext;
                for (int split = 0; split + 1 < kBuckets; ++split) {
                    if (left[split].count == 0u || right[split].count == 0u) continue;
                    const float c = regularization * (cost(left[split].m) + cost(right[split].m));
                    if (c < parent_cost && c < best) {
                        best = c;
                        axis = a;
                        middle = begin + left[split].count;
                    }
                }
            }
        }
        if (!(count > max_leaf || axis >= 0)) {
            Node &n = nodes[index];
            n.kind = ORACLE_TREE_LEAF;
            n.first = begin;
            n.count = count;
            for (uint32_t i = begin; i < end; ++i) {
                to_tree[emitters[i].id] = i;
                to_leaf[emitters[i].id] = index;
            }
            return index;
        }
        if (axis >= 0) {
            std::nth_element(emitters.begin() + begin, emitters.begin() + middle, emitters.begin() + end,
                             [axis](const Emitter &x, const Emitter &y) {
                                 const float xv = comp(x.centroid, axis), yv = comp(y.centroid, axis);
                                 return xv < yv || (xv == yv && x.id < y.id);
                             });
        }
        nodes[index].kind = ORACLE_TREE_INNER;
        const uint32_t l = build(begin, middle, index);
        const uint32_t r = build(middle, end, index);
        nodes[index].left = l;
        nodes[index].right = r;
        return index;
    }
};

void put_measure(const Measure &m, float *f, uint32_t *u) {
    const float v[11] = {m.bounds.lo.x, m.bounds.lo.y, m.bounds.lo.z, m.bounds.hi.x, m.bounds.hi.y, m.bounds.hi.z,
                         m.cone.axis.x, m.cone.axis.y, m.cone.axis.z, m.cone.theta_o, m.cone.theta_e};
    std::memcpy(f, v, sizeof(v));
    f[11] = m.energy;
    u[0] = m.bounds.empty ? 1u : 0u;
    u[1] = m.cone.empty ? 1u : 0u;
}

}  // namespace

extern "C" int oracle_light_tree_build(const oracle_tree_emitter *in, uint32_t count, uint32_t max_leaf,
                                       oracle_light_tree *out) {
    std::memset(out, 0, sizeof(*out));
    out->root = kInvalid;
    Builder b;
    b.max_leaf = std::max(max_leaf, 1u);
    b.emitters.reserve(count);
    std::vector<bool> seen(count, false);
    for (uint32_t i = 0; i < count; ++i) {
        const oracle_tree_emitter &s = in[i];
        Emitter e;
        e.m.bounds.lo = {s.bounds_min[0], s.bounds_min[1], s.bounds_min[2]};
        e.m.bounds.hi = {s.bounds_max[0], s.bounds_max[1], s.bounds_max[2]};
        e.m.bounds.empty = s.bounds_empty != 0;
        e.m.cone.axis = {s.axis[0], s.axis[1], s.axis[2]};
        e.m.cone.theta_o = s.theta_o;
        e.m.cone.theta_e = s.theta_e;
        e.m.cone.empty = s.orientation_empty != 0;
        e.m.energy = s.energy;
        e.m = sanitized(e.m);
        e.centroid = {s.centroid[0], s.centroid[1], s.centroid[2]};
        if (!finite3(e.centroid)) e.centroid = {};
        e.id = s.emitter_id;
        e.distant = s.distant != 0;
        if (e.id >= count || seen[e.id]) return -1;  // identities must be a dense permutation
        seen[e.id] = true;
        b.emitters.push_back(e);
    }
    if (count == 0u) return 0;
    b.to_tree.assign(count, kInvalid);
    b.to_leaf.assign(count, kInvalid);
    const uint32_t distant_begin = static_cast<uint32_t>(
        std::stable_partition(b.emitters.begin(), b.emitters.end(), [](const Emitter &e) { return !e.distant; }) -
        b.emitters.begin());
    const uint32_t root = b.add_node(kInvalid);
    b.nodes[root].kind = ORACLE_TREE_INNER;
    const uint32_t local = b.build(0u, distant_begin, root);
    const uint32_t distant = b.add_node(root);
    b.nodes[distant].kind = ORACLE_TREE_DISTANT;
    b.nodes[distant].first = distant_begin;
    b.nodes[distant].count = count - distant_begin;
    for (uint32_t i = distant_begin; i < count; ++i) {
        b.nodes[distant].m = merge(b.nodes[distant].m, b.emitters[i].m);
        b.to_tree[b.emitters[i].id] = i;
        b.to_leaf[b.emitters[i].id] = distant;
    }
    b.nodes[root].left = local;
    b.nodes[root].right = distant;
    b.nodes[root].m = merge(b.nodes[local].m, b.nodes[distant].m);

    out->root = root;
    out->node_count = static_cast<uint32_t>(b.nodes.size());
    out->emitter_count = count;
    out->nodes = static_cast<oracle_tree_node *>(std::calloc(out->node_count, sizeof(oracle_tree_node)));
    out->emitters = static_cast<oracle_tree_record *>(std::calloc(count, sizeof(oracle_tree_record)));
    out->emitter_to_tree = static_cast<uint32_t *>(std::calloc(count, sizeof(uint32_t)));
    out->emitter_to_leaf = static_cast<uint32_t *>(std::calloc(count, sizeof(uint32_t)));
    for (uint32_t i = 0; i < out->node_count; ++i) {
        const Node &n = b.nodes[i];
        oracle_tree_node &o = out->nodes[i];
        put_measure(n.m, o.measure, o.flags);
        o.parent = n.parent;
        o.left = n.left;
        o.right = n.right;
        o.first = n.first;
        o.count = n.count;
        o.kind = n.kind;
    }
    for (uint32_t i = 0; i < count; ++i) {
        const Emitter &e = b.emitters[i];
        oracle_tree_record &o = out->emitters[i];
        put_measure(e.m, o.measure, o.flags);
        o.centroid[0] = e.centroid.x;
        o.centroid[1] = e.centroid.y;
        o.centroid[2] = e.centroid.z;
        o.emitter_id = e.id;
        o.distant = e.distant ? 1u : 0u;
        out->emitter_to_tree[i] = b.to_tree[i];
        out->emitter_to_leaf[i] = b.to_leaf[i];
    }
    return 0;
}

extern "C" void oracle_light_tree_free(oracle_light_tree *tree) {
    if (!tree) return;
    std::free(tree->nodes);
    std::free(tree->emitters);
    std::free(tree->emitter_to_tree);
    std::free(tree->emitter_to_leaf);
    std::memset(tree, 0, sizeof(*tree));
}
