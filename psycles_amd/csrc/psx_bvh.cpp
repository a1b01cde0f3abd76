// psx_bvh.cpp - host builder of the two-level BVH4.
//
// Replaces the acceleration structure the reference obtains from
// LuisaCompute's Accel (device.create_mesh / accel.emplace_back /
// accel.build(), src/luisa/path_tracer_scene.cpp:1254-1255, 1675-1680;
// src/luisa/path_tracer_scene_upload.cpp:259-260).  A binned-SAH binary tree
// (32 bins, leaves of at most four items) is collapsed into 4-wide nodes of one
// 128-byte line each.  Every child box is padded outward so the traversal is
// conservative with respect to the exact triangle test: the BVH only prunes,
// it never decides which candidate wins.
#include "psx_bvh.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace psx {
namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Box &b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow_point(const float p[3]) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
    bool empty() const { return lo[0] > hi[0]; }
};

struct BinNode {
    Box box;
    uint32_t left = 0, right = 0;  // children (inner)
    uint32_t first = 0, count = 0; // item range (leaf) when count > 0
};

void pad(Box &b) {
    if (b.empty()) return;
    for (int a = 0; a < 3; ++a) {
        const float ext = b.hi[a] - b.lo[a];
        const float s = std::max(std::max(std::fabs(b.lo[a]), std::fabs(b.hi[a])), ext);
        const float p = s * 1.0e-5f + 1.0e-30f;
        b.lo[a] -= p;
        b.hi[a] += p;
    }
}

class BinaryBuilder {
  public:
    BinaryBuilder(const std::vector<BuildItem> &items, uint32_t max_leaf) : items_(items), leaf_size_(max_leaf) {
        order_.resize(items.size());
        centroids_.resize(items.size() * 3);
        for (size_t i = 0; i < items.size(); ++i) {
            order_[i] = static_cast<uint32_t>(i);
            for (int a = 0; a < 3; ++a) centroids_[3 * i + a] = 0.5f * (items[i].lo[a] + items[i].hi[a]);
        }
    }

    void build() {
        nodes_.clear();
        nodes_.reserve(items_.size() / 2 + 2);
        nodes_.push_back(BinNode{});
        struct Task {
            uint32_t node, begin, end, depth;
        };
        std::vector<Task> tasks;
        tasks.push_back({0u, 0u, static_cast<uint32_t>(items_.size()), 0u});
        while (!tasks.empty()) {
            const Task t = tasks.back();
            tasks.pop_back();
            Box box, cbox;
            for (uint32_t i = t.begin; i < t.end; ++i) {
                const BuildItem &it = items_[order_[i]];
                Box ib;
                std::memcpy(ib.lo, it.lo, sizeof(ib.lo));
                std::memcpy(ib.hi, it.hi, sizeof(ib.hi));
                box.grow(ib);
                cbox.grow_point(&centroids_[3 * order_[i]]);
            }
            nodes_[t.node].box = box;
            max_depth_ = std::max(max_depth_, t.depth);
            const uint32_t n = t.end - t.begin;
            if (n <= leaf_size_) {
                nodes_[t.node].first = t.begin;
                nodes_[t.node].count = n;
                continue;
            }
            const uint32_t mid = split(t.begin, t.end, cbox);
            const uint32_t left = static_cast<uint32_t>(nodes_.size());
            nodes_.push_back(BinNode{});
            nodes_.push_back(BinNode{});
            nodes_[t.node].left = left;
            nodes_[t.node].right = left + 1u;
            tasks.push_back({left + 1u, mid, t.end, t.depth + 1u});
            tasks.push_back({left, t.begin, mid, t.depth + 1u});
        }
    }

    const std::vector<BinNode> &nodes() const { return nodes_; }
    const std::vector<uint32_t> &order() const { return order_; }

  private:
    static constexpr int kBins = 32;

    uint32_t split(uint32_t begin, uint32_t end, const Box &cbox) {
        float best_cost = FLT_MAX;
        int best_axis = -1, best_bin = 0;
        for (int a = 0; a < 3; ++a) {
            const float extent = cbox.hi[a] - cbox.lo[a];
            if (!(extent > 0.0f)) continue;
            Box bins[kBins];
            uint32_t counts[kBins] = {0};
            const float scale = static_cast<float>(kBins) / extent;
            for (uint32_t i = begin; i < end; ++i) {
                const uint32_t id = order_[i];
                int b = static_cast<int>((centroids_[3 * id + a] - cbox.lo[a]) * scale);
                b = std::min(std::max(b, 0), kBins - 1);
                counts[b]++;
                Box ib;
                std::memcpy(ib.lo, items_[id].lo, sizeof(ib.lo));
                std::memcpy(ib.hi, items_[id].hi, sizeof(ib.hi));
                bins[b].grow(ib);
            }
            float right_area[kBins];
            uint32_t right_count[kBins];
            Box acc;
            uint32_t c = 0;
            for (int b = kBins - 1; b >= 0; --b) {
                acc.grow(bins[b]);
                c += counts[b];
                right_area[b] = acc.area();
                right_count[b] = c;
            }
            acc = Box{};
            c = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.grow(bins[b]);
                c += counts[b];
                if (c == 0 || right_count[b + 1] == 0) continue;
                const float cost = acc.area() * static_cast<float>(c) +
                                   right_area[b + 1] * static_cast<float>(right_count[b + 1]);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_bin = b;
                }
            }
        }
        uint32_t mid = begin + (end - begin) / 2u;
        if (best_axis >= 0) {
            const float extent = cbox.hi[best_axis] - cbox.lo[best_axis];
            const float scale = static_cast<float>(kBins) / extent;
            auto it = std::partition(order_.begin() + begin, order_.begin() + end, [&](uint32_t id) {
                int b = static_cast<int>((centroids_[3 * id + best_axis] - cbox.lo[best_axis]) * scale);
                b = std::min(std::max(b, 0), kBins - 1);
                return b <= best_bin;
            });
            mid = static_cast<uint32_t>(it - order_.begin());
        }
        if (mid == begin || mid == end) {
            // degenerate centroids: balanced median split on the widest axis
            int axis = 0;
            float w = -1.0f;
            for (int a = 0; a < 3; ++a)
                if (cbox.hi[a] - cbox.lo[a] > w) {
                    w = cbox.hi[a] - cbox.lo[a];
                    axis = a;
                }
            mid = begin + (end - begin) / 2u;
            std::nth_element(order_.begin() + begin, order_.begin() + mid, order_.begin() + end,
                             [&](uint32_t x, uint32_t y) {
                                 return centroids_[3 * x + axis] < centroids_[3 * y + axis];
                             });
        }
        return mid;
    }

    const std::vector<BuildItem> &items_;
    uint32_t leaf_size_;
    std::vector<uint32_t> order_;
    std::vector<float> centroids_;
    std::vector<BinNode> nodes_;
    uint32_t max_depth_ = 0;
};

}  // namespace

Bvh4 build_bvh4(const std::vector<BuildItem> &items, uint32_t node_base, uint32_t max_leaf, uint32_t width) {
    width = std::min(std::max(width, 2u), 4u);
    Bvh4 out;
    if (items.empty()) {
        Node4 n{};
        for (int i = 0; i < 4; ++i) {
            n.lo_x[i] = n.lo_y[i] = n.lo_z[i] = 1.0f;
            n.hi_x[i] = n.hi_y[i] = n.hi_z[i] = -1.0f;
            n.child[i] = 0u;
            n.count[i] = kEmptySlot;
        }
        out.nodes.push_back(n);
        out.depth = 1;
        return out;
    }
    BinaryBuilder builder(items, std::min(std::max(max_leaf, 1u), 4u));
    builder.build();
    const auto &bn = builder.nodes();
    out.order = builder.order();
    // collapse: each node absorbs up to `width` binary descendants, always
    // opening the inner child with the largest surface area; unused slots
    // of the 4-wide record stay empty.
    struct Task {
        uint32_t binary, slot_node, depth;
    };
    std::vector<Task> tasks;
    out.nodes.push_back(Node4{});
    tasks.push_back({0u, 0u, 1u});
    // a root that is itself a leaf becomes one leaf child
    while (!tasks.empty()) {
        const Task t = tasks.back();
        tasks.pop_back();
        out.depth = std::max(out.depth, t.depth);
        uint32_t kids[4];
        int k = 0;
        if (bn[t.binary].count > 0u) {
            kids[k++] = t.binary;
        } else {
            kids[k++] = bn[t.binary].left;
            kids[k++] = bn[t.binary].right;
            while (k < static_cast<int>(width)) {
                int best = -1;
                float best_area = -1.0f;
                for (int i = 0; i < k; ++i) {
                    if (bn[kids[i]].count > 0u) continue;
                    const float a = bn[kids[i]].box.area();
                    if (a > best_area) {
                        best_area = a;
                        best = i;
                    }
                }
                if (best < 0) break;
                const uint32_t open = kids[best];
                kids[best] = bn[open].left;
                kids[k++] = bn[open].right;
            }
        }
        Node4 node{};
        for (int i = 0; i < 4; ++i) {
            if (i >= k) {
                node.lo_x[i] = node.lo_y[i] = node.lo_z[i] = 1.0f;
                node.hi_x[i] = node.hi_y[i] = node.hi_z[i] = -1.0f;
                node.child[i] = 0u;
                node.count[i] = kEmptySlot;
                continue;
            }
            const BinNode &c = bn[kids[i]];
            Box b = c.box;
            pad(b);
            node.lo_x[i] = b.lo[0];
            node.lo_y[i] = b.lo[1];
            node.lo_z[i] = b.lo[2];
            node.hi_x[i] = b.hi[0];
            node.hi_y[i] = b.hi[1];
            node.hi_z[i] = b.hi[2];
            if (c.count > 0u) {
                node.child[i] = c.first;
                node.count[i] = c.count;
            } else {
                const uint32_t idx = static_cast<uint32_t>(out.nodes.size());
                out.nodes.push_back(Node4{});
                node.child[i] = node_base + idx;
                node.count[i] = 0u;
                tasks.push_back({kids[i], idx, t.depth + 1u});
            }
        }
        out.nodes[t.slot_node] = node;
    }
    // world/object bounds of the whole tree
    Box root = bn[0].box;
    pad(root);
    std::memcpy(out.lo, root.lo, sizeof(out.lo));
    std::memcpy(out.hi, root.hi, sizeof(out.hi));
    return out;
}

QNode4 quantize_node(const Node4 &n, bool tlas) {
    QNode4 q{};
    const float *lo[3] = {n.lo_x, n.lo_y, n.lo_z};
    const float *hi[3] = {n.hi_x, n.hi_y, n.hi_z};
    uint8_t *qlo[3] = {q.lo_x, q.lo_y, q.lo_z};
    uint8_t *qhi[3] = {q.hi_x, q.hi_y, q.hi_z};
    bool any = false;
    float nlo[3] = {0, 0, 0}, nhi[3] = {0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        if (n.count[i] == kEmptySlot) continue;
        for (int a = 0; a < 3; ++a) {
            nlo[a] = any ? std::min(nlo[a], lo[a][i]) : lo[a][i];
            nhi[a] = any ? std::max(nhi[a], hi[a][i]) : hi[a][i];
        }
        any = true;
    }
    for (int a = 0; a < 3; ++a) {
        // the smallest power-of-two step that spans the node in 255 steps
        q.origin[a] = nlo[a];
        const double extent = static_cast<double>(nhi[a]) - static_cast<double>(nlo[a]);
        int e = -126;
        if (extent > 0.0) {
            e = std::max(-126, static_cast<int>(std::ceil(std::log2(extent / 255.0))));
            while (e < 127 && std::ceil(extent / std::ldexp(1.0, e)) > 255.0) ++e;
        }
        q.exp[a] = static_cast<int8_t>(e);
        const double step = std::ldexp(1.0, e);
        for (int i = 0; i < 4; ++i) {
            if (n.count[i] == kEmptySlot) continue;
            // outward rounding, checked in exact arithmetic
            double l = std::floor((static_cast<double>(lo[a][i]) - nlo[a]) / step);
            double h = std::ceil((static_cast<double>(hi[a][i]) - nlo[a]) / step);
            l = std::min(std::max(l, 0.0), 255.0);
            h = std::min(std::max(h, 0.0), 255.0);
            while (l > 0.0 && static_cast<double>(nlo[a]) + l * step > lo[a][i]) l -= 1.0;
            while (h < 255.0 && static_cast<double>(nlo[a]) + h * step < hi[a][i]) h += 1.0;
            qlo[a][i] = static_cast<uint8_t>(l);
            qhi[a][i] = static_cast<uint8_t>(h);
        }
    }
    for (int i = 0; i < 4; ++i) {
        const uint32_t c = n.count[i], child = n.child[i];
        if (c == kEmptySlot)
            q.ref[i] = kEmptySlot;
        else if (c == 0u)
            q.ref[i] = child | (tlas ? kKindTlas : kKindBlas);
        else if (tlas)
            q.ref[i] = child | kKindInstance;
        else
            q.ref[i] = kKindLeaf | (child << 3) | (c - 1u);
    }
    return q;
}

}  // namespace psx
