// psx_bvh.h - host builder of BVH4 trees (see psx_bvh.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "psx_device.h"

namespace psx {

struct BuildItem {
    float lo[3];
    float hi[3];
};

struct Bvh4 {
    std::vector<Node4> nodes;     // node 0 is the root; inner child indices are
                                  // offset by the node_base passed to the build
    std::vector<uint32_t> order;  // leaf item order -> input item index
    float lo[3] = {0, 0, 0};
    float hi[3] = {0, 0, 0};
    uint32_t depth = 0;           // number of 4-wide levels
};

// max_leaf: largest number of items per leaf (1..4); width: children per
// node (2..4; the 4-wide record keeps unused slots empty).
Bvh4 build_bvh4(const std::vector<BuildItem> &items, uint32_t node_base, uint32_t max_leaf, uint32_t width = 4u);

// The traversal form of a node array: node i becomes QNode4 i with 8-bit
// outward-rounded child bounds and child references encoded as traversal
// entries.  `is_tlas(i)` tells whether node i belongs to the instance level
// (its leaves hold one instance index) or to a mesh level (its leaves hold
// count triangles from child[i]).
template <typename IsTlas>
std::vector<QNode4> quantize_nodes(const std::vector<Node4> &nodes, IsTlas is_tlas);

QNode4 quantize_node(const Node4 &n, bool tlas);

template <typename IsTlas>
std::vector<QNode4> quantize_nodes(const std::vector<Node4> &nodes, IsTlas is_tlas) {
    std::vector<QNode4> out(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) out[i] = quantize_node(nodes[i], is_tlas(static_cast<uint32_t>(i)));
    return out;
}

}  // namespace psx
