// psx_light_tree.h - host construction of the many-light tree.
//
// Cycles' light tree (src/sampling/light_tree.cpp:90-408): local emitters
// are split by 12-bucket surface-area-orientation cost, distant emitters form
// one branch at the root.  Node order (pre-order: node, left subtree, right
// subtree) and the emitter order inside leaves are part of the sampling
// contract, so the split uses std::nth_element with the reference's key.
#pragma once

#include <cstdint>
#include <vector>

#include "psx_device.h"

namespace psx {

struct TreeMeasureIn {
    float bounds_min[3], bounds_max[3];
    bool bounds_empty = true;
    float axis[3];
    float theta_o = 0.0f, theta_e = 0.0f;
    bool orientation_empty = true;
    float energy = 0.0f;
};

struct TreeEmitterIn {
    TreeMeasureIn measure;
    float centroid[3];
    uint32_t emitter_id = 0;
    bool distant = false;
};

struct LightTreeTables {
    std::vector<TreeNode> nodes;
    std::vector<TreeEmitter> emitters;  // tree order
    std::vector<uint2> mapping;         // by emitter id: (tree index, leaf node)
    uint32_t root = 0xffffffffu;
    bool usable = false;
};

// Returns false when emitter ids are not a dense permutation.
bool build_light_tree(const std::vector<TreeEmitterIn> &emitters, uint32_t max_leaf, LightTreeTables &out);

}  // namespace psx
