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

// max_leaf: largest number of items per leaf (1..4).
Bvh4 build_bvh4(const std::vector<BuildItem> &items, uint32_t node_base, uint32_t max_leaf);

}  // namespace psx
