// psx_light_tree.cpp - host construction of the many-light tree
// (build_cycles_light_tree, src/sampling/light_tree.cpp:90-408; packing as
// make_light_tree_scene_upload, src/luisa/path_tracer_light_tree_scene.cpp:
// 294-362).
#include "psx_light_tree.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <future>
#include <limits>

namespace psx {
namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kHalfPi = 0.5f * kPi;
constexpr float kTwoPi = 2.0f * kPi;
constexpr size_t kBucketCount = 12;
constexpr uint32_t kInvalid = 0xffffffffu;

struct V3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float norm(V3 a) { return std::sqrt(std::max(dot(a, a), 0.0f)); }
V3 unit_or(V3 v, V3 fallback) {
    const float n = norm(v);
    return n > 1.0e-20f ? v * (1.0f / n) : fallback;
}
float axis_value(V3 v, size_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
bool all_finite(V3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
V3 load(const float *p) { return {p[0], p[1], p[2]}; }

struct Box {
    V3 lo, hi;
    bool empty = true;
};
struct Orientation {
    V3 axis;
    float theta_o = 0.0f, theta_e = 0.0f;
    bool empty = true;
};
struct Measure {
    Box box;
    Orientation cone;
    float energy = 0.0f;
};
struct Item {
    Measure m;
    V3 c;
    uint32_t id = 0;
    bool distant = false;
};
struct BuildNode {
    Measure m;
    uint32_t parent = kInvalid, left = kInvalid, right = kInvalid, first = 0, count = 0;
    uint32_t kind = kTreeLeaf;
};

Box unite(const Box &a, const Box &b) {
    if (a.empty) return b;
    if (b.empty) return a;
    Box r;
    r.lo = {std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z)};
    r.hi = {std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z)};
    r.empty = false;
    return r;
}

// smallest cone holding both (wider cone rotated towards the narrower one)
Orientation unite(const Orientation &a, const Orientation &b) {
    if (a.empty) return b;
    if (b.empty) return a;
    const bool b_wider = b.theta_o > a.theta_o;
    const Orientation &wide = b_wider ? b : a;
    const Orientation &narrow = b_wider ? a : b;
    const float cosine = std::clamp(dot(wide.axis, narrow.axis), -1.0f, 1.0f);
    const float between = std::acos(cosine);
    const float theta_e = std::max(a.theta_e, b.theta_e);
    Orientation r;
    r.empty = false;
    r.theta_e = theta_e;
    if (wide.theta_o + 5.0e-4f >= std::min(kPi, between + narrow.theta_o)) {
        r.axis = wide.axis;
        r.theta_o = wide.theta_o;
        return r;
    }
    const float theta_o = 0.5f * (between + wide.theta_o + narrow.theta_o);
    if (theta_o >= kPi) {
        r.axis = wide.axis;
        r.theta_o = kPi;
        return r;
    }
    const float turn = theta_o - wide.theta_o;
    V3 side = narrow.axis - wide.axis * cosine;
    if (cosine < -0.9995f) {
        const V3 w = wide.axis;
        const V3 c = std::abs(w.x) < 0.9f ? V3{1.0f, 0.0f, 0.0f} : V3{0.0f, 1.0f, 0.0f};
        side = unit_or({w.y * c.z - w.z * c.y, w.z * c.x - w.x * c.z, w.x * c.y - w.y * c.x}, {0.0f, 0.0f, 1.0f});
    } else {
        side = unit_or(side, {0.0f, 0.0f, 1.0f});
    }
    r.axis = unit_or(wide.axis * std::cos(turn) + side * std::sin(turn), wide.axis);
    r.theta_o = theta_o;
    return r;
}

Measure unite(const Measure &a, const Measure &b) {
    if (!(a.energy > 0.0f)) return b;
    if (!(b.energy > 0.0f)) return a;
    return Measure{unite(a.box, b.box), unite(a.cone, b.cone), a.energy + b.energy};
}

float saoh_cost(const Measure &m) {
    if (!(m.energy > 0.0f) || m.cone.empty) return 0.0f;
    float area = 0.0f;
    if (!m.box.empty) {
        const V3 e = m.box.hi - m.box.lo;
        area = 2.0f * (e.x * e.y + e.x * e.z + e.y * e.z);
        if (area == 0.0f) area = norm(e);
    }
    const float theta_w = std::min(kPi, m.cone.theta_o + m.cone.theta_e);
    const float co = std::cos(m.cone.theta_o), so = std::sin(m.cone.theta_o);
    const float solid = kTwoPi * (1.0f - co) + kHalfPi * (2.0f * theta_w * so -
                                                            std::cos(m.cone.theta_o - 2.0f * theta_w) -
                                                            2.0f * m.cone.theta_o * so + co);
    return m.energy * area * solid;
}

Measure sanitize(Measure m) {
    if (!std::isfinite(m.energy) || m.energy <= 0.0f) return Measure{};
    if (!m.box.empty && (!all_finite(m.box.lo) || !all_finite(m.box.hi))) m.box = Box{};
    if (!m.cone.empty) {
        m.cone.axis = unit_or(m.cone.axis, {0.0f, 0.0f, 1.0f});
        m.cone.theta_o = std::clamp(m.cone.theta_o, 0.0f, kPi);
        m.cone.theta_e = std::clamp(m.cone.theta_e, 0.0f, kPi);
    }
    return m;
}

struct Slot {
    Measure m;
    uint32_t count = 0;
};

// Split costs along one axis: bucket measures united in item order, one cost
// per bucket boundary (infinite where a side is empty).
struct AxisCosts {
    std::array<float, kBucketCount - 1> cost;
    std::array<uint32_t, kBucketCount - 1> below;
};

// Nodes with at least kParallelItems emitters unite their measure and
// evaluate their three axes on separate threads.  Down to kParallelDepth
// levels, a node whose children both hold at least kSubtreeItems emitters
// builds its right subtree on another thread while the left one is built
// here.  Every float operation happens in the same order as in a serial
// build, and subtrees are numbered as the serial pre-order numbers them, so
// the tables are identical.
constexpr uint32_t kParallelItems = 4096u;
constexpr uint32_t kParallelDepth = 5u;
constexpr uint32_t kSubtreeItems = 512u;

class TreeBuilder {
  public:
    // shared by the builders of disjoint subtrees: items [begin, end) of a
    // subtree, and the entries of tree_of / leaf_of for their ids
    std::vector<Item> *items = nullptr;
    std::vector<uint32_t> *tree_of = nullptr, *leaf_of = nullptr;
    std::vector<BuildNode> nodes;  // this builder's nodes, numbered from 0
    uint32_t leaf_size = 8;

    uint32_t new_node(uint32_t parent) {
        BuildNode n;
        n.parent = parent;
        nodes.push_back(n);
        return static_cast<uint32_t>(nodes.size()) - 1u;
    }

    AxisCosts axis_costs(uint32_t begin, uint32_t end, size_t a, const Box &centres, float widest) const {
        AxisCosts r;
        r.cost.fill(std::numeric_limits<float>::infinity());
        r.below.fill(0u);
        const V3 extent = centres.hi - centres.lo;
        const float span = axis_value(extent, a);
        if (!(span > 0.0f)) return r;
        const std::vector<Item> &it = *items;
        std::array<Slot, kBucketCount> slots{};
        for (uint32_t i = begin; i < end; ++i) {
            const float rel = (axis_value(it[i].c, a) - axis_value(centres.lo, a)) / span;
            const size_t s = std::min<size_t>(
                static_cast<size_t>(std::max(rel, 0.0f) * static_cast<float>(kBucketCount)), kBucketCount - 1);
            slots[s].m = unite(slots[s].m, it[i].m);
            ++slots[s].count;
        }
        std::array<Slot, kBucketCount - 1> below{}, above{};
        below[0] = slots[0];
        for (size_t s = 1; s + 1 < kBucketCount; ++s)
            below[s] = Slot{unite(below[s - 1].m, slots[s].m), below[s - 1].count + slots[s].count};
        above[kBucketCount - 2] = slots[kBucketCount - 1];
        for (size_t s = kBucketCount - 2; s-- > 0;)
            above[s] = Slot{unite(above[s + 1].m, slots[s + 1].m), above[s + 1].count + slots[s + 1].count};
        const float stretch = widest / span;
        for (size_t s = 0; s + 1 < kBucketCount; ++s) {
            if (below[s].count == 0u || above[s].count == 0u) continue;
            r.cost[s] = stretch * (saoh_cost(below[s].m) + saoh_cost(above[s].m));
            r.below[s] = below[s].count;
        }
        return r;
    }

    uint32_t local(uint32_t begin, uint32_t end, uint32_t parent, uint32_t depth) {
        std::vector<Item> &it = *items;
        const uint32_t self = new_node(parent);
        const uint32_t n = end - begin;
        if (n == 0u) return self;
        Box centres;
        for (uint32_t i = begin; i < end; ++i) centres = unite(centres, Box{it[i].c, it[i].c, false});
        const V3 extent = centres.hi - centres.lo;
        const float widest = std::max({extent.x, extent.y, extent.z});
        auto unite_all = [&] {
            Measure all;
            for (uint32_t i = begin; i < end; ++i) all = unite(all, it[i].m);
            return all;
        };
        Measure all;
        std::array<AxisCosts, 3> axes;
        bool costed = false;
        if (n >= kParallelItems && widest > 0.0f) {
            // the axes' costs do not depend on the node's measure: all four
            // passes at once, the parent-cost test afterwards
            auto m = std::async(std::launch::async, unite_all);
            auto y = std::async(std::launch::async, [&] { return axis_costs(begin, end, 1, centres, widest); });
            auto z = std::async(std::launch::async, [&] { return axis_costs(begin, end, 2, centres, widest); });
            axes[0] = axis_costs(begin, end, 0, centres, widest);
            axes[1] = y.get();
            axes[2] = z.get();
            all = m.get();
            costed = true;
        } else {
            all = unite_all();
        }
        nodes[self].m = all;
        size_t split_axis = std::numeric_limits<size_t>::max();
        uint32_t middle = begin + n / 2u;
        const float parent_cost = saoh_cost(all);
        if (n > 1u && parent_cost > 0.0f && widest > 0.0f) {
            if (!costed)
                for (size_t a = 0; a < 3; ++a) axes[a] = axis_costs(begin, end, a, centres, widest);
            // axes and bucket boundaries in order, each replacing the best
            // only when strictly cheaper
            float best = std::numeric_limits<float>::infinity();
            for (size_t a = 0; a < 3; ++a) {
                for (size_t b = 0; b + 1 < kBucketCount; ++b) {
                    const float c = axes[a].cost[b];
                    if (c < parent_cost && c < best) {
                        best = c;
                        split_axis = a;
                        middle = begin + axes[a].below[b];
                    }
                }
            }
        }
        const bool have_split = split_axis != std::numeric_limits<size_t>::max();
        if (!(n > leaf_size || have_split)) {
            BuildNode &leaf = nodes[self];
            leaf.kind = kTreeLeaf;
            leaf.first = begin;
            leaf.count = n;
            for (uint32_t i = begin; i < end; ++i) {
                (*tree_of)[it[i].id] = i;
                (*leaf_of)[it[i].id] = self;
            }
            return self;
        }
        if (have_split) {
            const size_t a = split_axis;
            std::nth_element(it.begin() + begin, it.begin() + middle, it.begin() + end,
                             [a](const Item &p, const Item &q) {
                                 const float pv = axis_value(p.c, a), qv = axis_value(q.c, a);
                                 return pv < qv || (pv == qv && p.id < q.id);
                             });
        }
        nodes[self].kind = kTreeInner;
        if (depth < kParallelDepth && end - middle >= kSubtreeItems && middle - begin >= kSubtreeItems) {
            // the right subtree on its own builder, numbered from 0, then
            // appended after the left subtree with its indices shifted
            TreeBuilder right;
            right.items = items;
            right.tree_of = tree_of;
            right.leaf_of = leaf_of;
            right.leaf_size = leaf_size;
            auto done = std::async(std::launch::async, [&] { right.local(middle, end, kInvalid, depth + 1u); });
            const uint32_t l = local(begin, middle, self, depth + 1u);
            done.get();
            const uint32_t offset = static_cast<uint32_t>(nodes.size());
            auto shift = [offset](uint32_t v) { return v == kInvalid ? kInvalid : v + offset; };
            for (BuildNode node : right.nodes) {
                node.parent = node.parent == kInvalid ? self : node.parent + offset;
                node.left = shift(node.left);
                node.right = shift(node.right);
                nodes.push_back(node);
            }
            for (uint32_t i = middle; i < end; ++i) (*leaf_of)[it[i].id] += offset;
            nodes[self].left = l;
            nodes[self].right = offset;
            return self;
        }
        const uint32_t l = local(begin, middle, self, depth + 1u);
        const uint32_t r = local(middle, end, self, depth + 1u);
        nodes[self].left = l;
        nodes[self].right = r;
        return self;
    }
};

uint32_t measure_flags(const Measure &m, bool distant) {
    return (m.box.empty ? 0u : kTreeHasBounds) | (m.cone.empty ? 0u : kTreeHasOrientation) |
           (distant ? kTreeDistant : 0u);
}

void pack(const Measure &m, float4 &min_energy, float4 &max_theta_o, float4 &axis_theta_e, float4 &trig) {
    min_energy = make_float4(m.box.lo.x, m.box.lo.y, m.box.lo.z, m.energy);
    max_theta_o = make_float4(m.box.hi.x, m.box.hi.y, m.box.hi.z, m.cone.theta_o);
    axis_theta_e = make_float4(m.cone.axis.x, m.cone.axis.y, m.cone.axis.z, m.cone.theta_e);
    // the angle terms of the importance bound depend on the measure only
    const float o = m.cone.theta_o, e = m.cone.theta_e;
    trig = make_float4(::cosf(o), ::sinf(o), ::cosf(o + e), ::cosf(e - o));
}

}  // namespace

bool build_light_tree(const std::vector<TreeEmitterIn> &emitters, uint32_t max_leaf, LightTreeTables &out) {
    out = LightTreeTables{};
    const uint32_t count = static_cast<uint32_t>(emitters.size());
    std::vector<Item> items;
    std::vector<uint32_t> tree_of, leaf_of;
    TreeBuilder b;
    b.items = &items;
    b.tree_of = &tree_of;
    b.leaf_of = &leaf_of;
    b.leaf_size = std::max(max_leaf, 1u);
    std::vector<bool> seen(count, false);
    for (const TreeEmitterIn &e : emitters) {
        if (e.emitter_id >= count || seen[e.emitter_id]) return false;
        seen[e.emitter_id] = true;
        Item it;
        it.m.box = Box{load(e.measure.bounds_min), load(e.measure.bounds_max), e.measure.bounds_empty};
        it.m.cone = Orientation{load(e.measure.axis), e.measure.theta_o, e.measure.theta_e,
                                e.measure.orientation_empty};
        it.m.energy = e.measure.energy;
        it.m = sanitize(it.m);
        it.c = load(e.centroid);
        if (!all_finite(it.c)) it.c = V3{};
        it.id = e.emitter_id;
        it.distant = e.distant;
        items.push_back(it);
    }
    if (count == 0u) return true;
    tree_of.assign(count, kInvalid);
    leaf_of.assign(count, kInvalid);
    const uint32_t distant_begin = static_cast<uint32_t>(
        std::stable_partition(items.begin(), items.end(), [](const Item &it) { return !it.distant; }) -
        items.begin());
    const uint32_t root = b.new_node(kInvalid);
    b.nodes[root].kind = kTreeInner;
    const uint32_t near = b.local(0u, distant_begin, root, 0u);
    const uint32_t far = b.new_node(root);
    b.nodes[far].kind = kTreeDistantNode;
    b.nodes[far].first = distant_begin;
    b.nodes[far].count = count - distant_begin;
    for (uint32_t i = distant_begin; i < count; ++i) {
        b.nodes[far].m = unite(b.nodes[far].m, items[i].m);
        tree_of[items[i].id] = i;
        leaf_of[items[i].id] = far;
    }
    b.nodes[root].left = near;
    b.nodes[root].right = far;
    b.nodes[root].m = unite(b.nodes[near].m, b.nodes[far].m);

    out.root = root;
    out.nodes.resize(b.nodes.size());
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        const BuildNode &n = b.nodes[i];
        TreeNode &g = out.nodes[i];
        pack(n.m, g.bounds_min_energy, g.bounds_max_theta_o, g.cone_axis_theta_e, g.trig);
        g.topology = make_uint4(n.parent, n.left, n.right, n.kind);
        g.emitters = make_uint4(n.first, n.count, measure_flags(n.m, n.kind == kTreeDistantNode), 0u);
    }
    out.emitters.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        const Item &it = items[i];
        TreeEmitter &g = out.emitters[i];
        pack(it.m, g.bounds_min_energy, g.bounds_max_theta_o, g.cone_axis_theta_e, g.trig);
        g.identity = make_uint4(it.id, measure_flags(it.m, it.distant), 0u, 0u);
    }
    out.mapping.resize(count);
    for (uint32_t id = 0; id < count; ++id) out.mapping[id] = make_uint2(tree_of[id], leaf_of[id]);
    out.usable = b.nodes[root].m.energy > 0.0f;
    return true;
}

}  // namespace psx
