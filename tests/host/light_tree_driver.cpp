// light_tree_driver.cpp - runs the library's host light-tree builder
// (psycles_amd/csrc/psx_light_tree.cpp) on emitters read from a file and
// writes its tables, so CPU tests can compare them with the oracle's build.
//
// input:  u32 count, u32 max_leaf, count x 19-word emitter records
//         (bounds_min[3], bounds_max[3], i32 bounds_empty, axis[3], theta_o,
//         theta_e, i32 orientation_empty, energy, centroid[3], u32 id,
//         i32 distant)
// output: u32 ok, u32 root, u32 node_count, node_count x 24 words
//         (TreeNode), count x 20 words (TreeEmitter), count x 2 words
//         (mapping)
#include <cstdio>
#include <cstring>
#include <vector>

#include "psx_light_tree.h"

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    uint32_t head[2];
    if (std::fread(head, 4, 2, in) != 2) return 2;
    std::vector<psx::TreeEmitterIn> emitters(head[0]);
    for (psx::TreeEmitterIn &e : emitters) {
        uint32_t w[19];
        if (std::fread(w, 4, 19, in) != 19) return 2;
        auto f = [&](int i) {
            float v;
            std::memcpy(&v, &w[i], 4);
            return v;
        };
        for (int k = 0; k < 3; ++k) {
            e.measure.bounds_min[k] = f(k);
            e.measure.bounds_max[k] = f(3 + k);
            e.measure.axis[k] = f(7 + k);
            e.centroid[k] = f(14 + k);
        }
        e.measure.bounds_empty = w[6] != 0u;
        e.measure.theta_o = f(10);
        e.measure.theta_e = f(11);
        e.measure.orientation_empty = w[12] != 0u;
        e.measure.energy = f(13);
        e.emitter_id = w[17];
        e.distant = w[18] != 0u;
    }
    std::fclose(in);
    psx::LightTreeTables t;
    const uint32_t ok = psx::build_light_tree(emitters, head[1], t) ? 1u : 0u;
    FILE *out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    const uint32_t info[3] = {ok, t.root, static_cast<uint32_t>(t.nodes.size())};
    std::fwrite(info, 4, 3, out);
    std::fwrite(t.nodes.data(), sizeof(psx::TreeNode), t.nodes.size(), out);
    std::fwrite(t.emitters.data(), sizeof(psx::TreeEmitter), t.emitters.size(), out);
    std::fwrite(t.mapping.data(), sizeof(uint2), t.mapping.size(), out);
    std::fclose(out);
    return 0;
}
