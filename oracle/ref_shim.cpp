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
#include <psycles/sampling/light_distribution.h>
#include <psycles/sampling/pixel_filter.h>
#include <psycles/sampling/tabulated_sobol.h>

#include <cstdint>
#include <cstring>
#include <span>

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

}  // extern "C"
