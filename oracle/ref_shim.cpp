// ref_shim.cpp - C bindings over the reference's own Luisa-free sampling
// sources, compiled in place from /root/reference by oracle/Makefile into
// oracle/_ref/libpsycles_ref.so.  TEST INFRASTRUCTURE ONLY: used to validate
// the C restatement in psx_oracle.c, never by the product.
//
// Bound reference functions:
//   tabulated_sobol::generate_table / sequence_size_for_samples
//       (src/sampling/tabulated_sobol.cpp:156-199)
//   sampling::make_pixel_filter_table (src/sampling/pixel_filter.cpp:30-86)
//   sampling::build_cycles_light_distribution
//       (src/sampling/light_distribution.cpp:25-138)
//   sampling::build_cycles_light_tree (src/sampling/light_tree.cpp:403-408),
//       emitted in the oracle_light_tree layout of psx_oracle.h
#include <psycles/sampling/light_distribution.h>
#include <psycles/sampling/light_tree.h>
#include <psycles/sampling/pixel_filter.h>
#include <psycles/sampling/tabulated_sobol.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <span>
#include <vector>

#include "psx_oracle.h"

namespace {

void put_measure(const psycles::sampling::LightTreeMeasure &m, float *f, std::uint32_t *u) {
    const float v[12] = {m.bounds.minimum.x, m.bounds.minimum.y, m.bounds.minimum.z, m.bounds.maximum.x,
                         m.bounds.maximum.y, m.bounds.maximum.z, m.orientation.axis.x, m.orientation.axis.y,
                         m.orientation.axis.z, m.orientation.theta_o, m.orientation.theta_e, m.energy};
    std::memcpy(f, v, sizeof(v));
    u[0] = m.bounds.empty ? 1u : 0u;
    u[1] = m.orientation.empty ? 1u : 0u;
}

}  // namespace

extern "C" {

std::uint32_t ref_sequence_size_for_samples(std::uint32_t count) {
    return psycles::sampling::tabulated_sobol::sequence_size_for_samples(count);
}

int ref_sobol_table(std::uint32_t sequence_size, float *out) {
    try {
        const auto table =
            psycles::sampling::tabulated_sobol::generate_table(sequence_size);
        std::memcpy(out, table.data(), table.size() * sizeof(float) * 4u);
        return 0;
    } catch (...) {
        return -1;
    }
}

void ref_pixel_filter_table(std::uint32_t filter, float width, float *out) {
    const auto table = psycles::sampling::make_pixel_filter_table(
        static_cast<psycles::contract::PixelFilter>(filter), width);
    std::memcpy(out, table.data(), table.size() * sizeof(float));
}

std::uint32_t ref_light_distribution(const float *areas, std::uint32_t count,
                                     std::uint32_t analytic_count,
                                     int include_environment,
                                     std::uint32_t *entries,
                                     float *triangle_area_pdf,
                                     float *light_selection_pdf) {
    const auto d = psycles::sampling::build_cycles_light_distribution(
        std::span<const float>{areas, count}, analytic_count,
        include_environment != 0);
    for (std::size_t i = 0; i < d.entries.size(); ++i) {
        std::memcpy(&entries[4 * i + 0], &d.entries[i].cumulative, 4);
        std::memcpy(&entries[4 * i + 1], &d.entries[i].selection_pdf, 4);
        entries[4 * i + 2] = static_cast<std::uint32_t>(d.entries[i].kind);
        entries[4 * i + 3] = d.entries[i].index;
    }
    *triangle_area_pdf = d.triangle_area_pdf;
    *light_selection_pdf = d.light_selection_pdf;
    return d.emitter_count;
}

int ref_light_tree_build(const oracle_tree_emitter *in, std::uint32_t count, std::uint32_t max_leaf,
                         oracle_light_tree *out) {
    using namespace psycles::sampling;
    std::memset(out, 0, sizeof(*out));
    std::vector<LightTreeEmitter> emitters(count);
    for (std::uint32_t i = 0; i < count; ++i) {
        const oracle_tree_emitter &s = in[i];
        LightTreeEmitter &e = emitters[i];
        e.measure.bounds.minimum = {s.bounds_min[0], s.bounds_min[1], s.bounds_min[2]};
        e.measure.bounds.maximum = {s.bounds_max[0], s.bounds_max[1], s.bounds_max[2]};
        e.measure.bounds.empty = s.bounds_empty != 0;
        e.measure.orientation.axis = {s.axis[0], s.axis[1], s.axis[2]};
        e.measure.orientation.theta_o = s.theta_o;
        e.measure.orientation.theta_e = s.theta_e;
        e.measure.orientation.empty = s.orientation_empty != 0;
        e.measure.energy = s.energy;
        e.centroid = {s.centroid[0], s.centroid[1], s.centroid[2]};
        e.emitter_id = s.emitter_id;
        e.distant = s.distant != 0;
    }
    CyclesLightTree tree;
    try {
        tree = build_cycles_light_tree(emitters, max_leaf);
    } catch (...) {
        return -1;
    }
    out->root = tree.root;
    out->node_count = static_cast<std::uint32_t>(tree.nodes.size());
    out->emitter_count = static_cast<std::uint32_t>(tree.emitters.size());
    out->nodes = static_cast<oracle_tree_node *>(std::calloc(out->node_count + 1u, sizeof(oracle_tree_node)));
    out->emitters =
        static_cast<oracle_tree_record *>(std::calloc(out->emitter_count + 1u, sizeof(oracle_tree_record)));
    out->emitter_to_tree = static_cast<std::uint32_t *>(std::calloc(out->emitter_count + 1u, 4u));
    out->emitter_to_leaf = static_cast<std::uint32_t *>(std::calloc(out->emitter_count + 1u, 4u));
    for (std::uint32_t i = 0; i < out->node_count; ++i) {
        const auto &n = tree.nodes[i];
        auto &o = out->nodes[i];
        put_measure(n.measure, o.measure, o.flags);
        o.parent = n.parent;
        o.left = n.left;
        o.right = n.right;
        o.first = n.first_emitter;
        o.count = n.emitter_count;
        o.kind = static_cast<std::uint32_t>(n.kind);
    }
    for (std::uint32_t i = 0; i < out->emitter_count; ++i) {
        const auto &e = tree.emitters[i];
        auto &o = out->emitters[i];
        put_measure(e.measure, o.measure, o.flags);
        o.centroid[0] = e.centroid.x;
        o.centroid[1] = e.centroid.y;
        o.centroid[2] = e.centroid.z;
        o.emitter_id = e.emitter_id;
        o.distant = e.distant ? 1u : 0u;
        out->emitter_to_tree[i] = tree.emitter_to_tree[i];
        out->emitter_to_leaf[i] = tree.emitter_to_leaf[i];
    }
    return 0;
}

void ref_light_tree_free(oracle_light_tree *tree) {
    std::free(tree->nodes);
    std::free(tree->emitters);
    std::free(tree->emitter_to_tree);
    std::free(tree->emitter_to_leaf);
    std::memset(tree, 0, sizeof(*tree));
}

}  // extern "C"
