/*
 * psx_oracle.h - CPU restatement of Psycles' device render path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the HIP render path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * Plain C99, built with -O2 -ffp-contract=off and without fast-math so that
 * every floating-point operation happens in the order written.  Each function
 * cites the reference file it restates.  Scene and settings use the public
 * ABI structs from include/psx.h.
 */
#ifndef PSX_ORACLE_H
#define PSX_ORACLE_H

#include <stdint.h>

#include "../include/psx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- tabulated Sobol (src/sampling/tabulated_sobol.cpp) ---------------- */
uint32_t oracle_sequence_size_for_samples(uint32_t sample_count);
/* out: 4 * 256 * sequence_size floats; returns 0 or -1 for a bad size. */
int oracle_sobol_table(uint32_t sequence_size, float *out);

/* --- device sampler (include/psycles/luisa/cycles_sampler.h) ----------- */
uint32_t oracle_pixel_hash(uint32_t x, uint32_t y, uint32_t seed);
uint32_t oracle_shuffled_sample_index(uint32_t sample, uint32_t dimension,
                                      uint32_t seed, uint32_t sequence_size);

/* --- pixel filter (src/sampling/pixel_filter.cpp) ----------------------- */
void oracle_pixel_filter_table(uint32_t filter, float width, float *out1024);
float oracle_pixel_filter_sample(const float *table, float random);

/* --- sample mappings (include/psycles/luisa/cycles_sample_mapping.h) ---- */
void oracle_sample_uniform_disk(const float random[2], float out[2]);
/* out: direction xyz, pdf */
void oracle_sample_cosine_hemisphere(const float normal[3],
                                     const float random[2], float out[4]);

/* --- camera (src/luisa/path_tracer_camera.cpp) -------------------------- */
void oracle_camera_clip_range(float near_clip, float far_clip, float cosine,
                              float out[2]);
void oracle_affine_point(const float m[16], const float p[3], float out[3]);

/* --- path state (include/psycles/luisa/cycles_path_state.h) ------------- */
typedef struct oracle_path_state {
    uint32_t flag, visibility, bounce, diffuse_bounce, glossy_bounce,
        transmission_bounce, transparent_bounce, rng_offset;
} oracle_path_state;
typedef struct oracle_path_limits {
    uint32_t maximum, maximum_diffuse, maximum_glossy, maximum_transmission,
        maximum_transparent;
} oracle_path_limits;
void oracle_initial_state(oracle_path_state *state);
void oracle_next_surface(const oracle_path_state *before, uint32_t label,
                         uint32_t runtime_flag,
                         const oracle_path_limits *limits,
                         oracle_path_state *after);
uint32_t oracle_contract_visibility(uint32_t cycles_visibility);
float oracle_continuation_probability(uint32_t flag, uint32_t bounce,
                                      uint32_t transparent_bounce,
                                      uint32_t minimum,
                                      uint32_t transparent_minimum,
                                      const float throughput[3]);

/* --- light distribution (src/sampling/light_distribution.cpp) ----------- */
/* entries: 4 words per entry (cumulative f32, selection_pdf f32, kind u32,
 * index u32), emitter_count + 1 entries.  Returns emitter_count. */
uint32_t oracle_light_distribution(const float *areas, uint32_t triangle_count,
                                   uint32_t analytic_count,
                                   int include_environment,
                                   uint32_t *entries,
                                   float *triangle_area_pdf,
                                   float *light_selection_pdf);

/* --- triangle intersection (cycles_triangle_intersection_component.cpp) -- */
/* out: valid, distance, u, v.  world_to_object may be NULL (identity). */
int oracle_pluecker_intersect(const float origin[3], const float direction[3],
                              float tmin, float tmax, const float p0[3],
                              const float p1[3], const float p2[3],
                              float out[3]);

/* --- analytic lamps (include/psycles/luisa/analytic_light_sampling.h) ---- */
/* out: direction[3], pdf, position[3], distance, normal[3], eval factor,
 * valid (1/0). */
int oracle_sample_lamp(const psx_light *light, const float reference[3],
                       const float shading_normal[3], int has_transmission,
                       const float random[2], float out[13]);
/* Forward hit of a point or spot lamp: out = valid, distance, pdf, eval. */
int oracle_intersect_lamp(const psx_light *light, const float origin[3],
                          const float direction[3], float tmin, float tmax,
                          const float previous_normal[3], int had_transmission,
                          float out[4]);

/* --- light tree (src/sampling/light_tree.cpp, built by
 *     psx_oracle_tree.cpp) ------------------------------------------------- */
enum { ORACLE_TREE_INNER = 0, ORACLE_TREE_LEAF = 1, ORACLE_TREE_DISTANT = 2 };

/* LightTreeEmitter input (include/psycles/sampling/light_tree.h:15-40). */
typedef struct oracle_tree_emitter {
    float bounds_min[3], bounds_max[3];
    int bounds_empty;
    float axis[3], theta_o, theta_e;
    int orientation_empty;
    float energy;
    float centroid[3];
    uint32_t emitter_id;
    int distant;
} oracle_tree_emitter;

/* measure[12] = bounds min (3), bounds max (3), axis (3), theta_o, theta_e,
 * energy; flags[2] = bounds empty, orientation empty. */
typedef struct oracle_tree_node {
    float measure[12];
    uint32_t flags[2];
    uint32_t parent, left, right, first, count, kind;
} oracle_tree_node;

typedef struct oracle_tree_record {
    float measure[12];
    uint32_t flags[2];
    float centroid[3];
    uint32_t emitter_id, distant;
} oracle_tree_record;

typedef struct oracle_light_tree {
    oracle_tree_node *nodes;
    oracle_tree_record *emitters;       /* tree (reordered) emitter order */
    uint32_t *emitter_to_tree;          /* indexed by emitter id          */
    uint32_t *emitter_to_leaf;
    uint32_t node_count, emitter_count, root;
} oracle_light_tree;

/* Returns 0, or -1 when identities are not a dense permutation. */
int oracle_light_tree_build(const oracle_tree_emitter *emitters, uint32_t count,
                            uint32_t max_emitters_per_leaf,
                            oracle_light_tree *out);
void oracle_light_tree_free(oracle_light_tree *tree);


/* --- render ------------------------------------------------------------- */
typedef struct oracle_stats {
    uint64_t rays;
    uint64_t node_visits;      /* binary-SAH nodes visited, all rays       */
    uint64_t triangle_tests;   /* triangles tested, all rays               */
    uint64_t samples;
} oracle_stats;

typedef struct oracle_scene oracle_scene;

/* Scene light tree (built by oracle_scene_create): one selection at (point,
 * normal) for random number u, with the sample's pdf and the reverse pdf of
 * the chosen emitter; and the reverse pdf of any emitter id. */
int oracle_light_tree_query(const oracle_scene *scene, const float point[3],
                            const float normal[3], int has_transmission, float u,
                            uint32_t *emitter_id, float *pdf_sample,
                            float *pdf_reverse);
float oracle_light_tree_pdf(const oracle_scene *scene, uint32_t emitter_id,
                            const float point[3], const float normal[3],
                            int has_transmission);
uint32_t oracle_light_tree_emitters(const oracle_scene *scene);
/* The same descent and reverse pdf on a tree built by oracle_light_tree_build
 * (emitter ids as given to the build). */
int oracle_light_tree_sample(const oracle_light_tree *tree, const float point[3], const float normal[3],
                             int has_transmission, float u, uint32_t *emitter_id, float *pdf_sample,
                             float *pdf_reverse);
float oracle_light_tree_reverse_pdf(const oracle_light_tree *tree, uint32_t emitter_id, const float point[3],
                                    const float normal[3], int has_transmission);

/* Build the oracle's own scene tables (binary SAH BVH with 16 bins and
 * leaves of at most 4 triangles, light distribution).  `brute_force`
 * replaces the BVH with an O(T) loop over every triangle. */
oracle_scene *oracle_scene_create(const psx_scene_desc *desc, int brute_force);
void oracle_scene_destroy(oracle_scene *scene);
int oracle_scene_update_parameters(oracle_scene *scene, uint32_t material,
                                   const float *parameters, uint32_t count);
uint32_t oracle_scene_emissive_triangles(const oracle_scene *scene);
/* Flat light distribution entries (count returned; at most `capacity`
 * written); kind 0 triangle, 1 lamp. */
uint32_t oracle_scene_distribution(const oracle_scene *scene, float *cumulative, float *selection_pdf,
                                   uint32_t *kind, uint32_t *index, uint32_t capacity);
/* A non-owning view of the scene's light tree: 0 when a usable tree exists,
 * 1 when the scene has none.  Never pass the view to oracle_light_tree_free. */
int oracle_scene_light_tree(const oracle_scene *scene, oracle_light_tree *view);

/* Accumulate samples [first, first + count) of a `total`-sample render into
 * the window film (row-major, same layout as psx_session_read_film).  The
 * film arrays must hold the previous sums (zero for a fresh render).
 * `threads` <= 0 uses every online CPU.  Returns 0 on success. */
int oracle_render(const oracle_scene *scene,
                  const psx_render_settings *settings, uint32_t first,
                  uint32_t count, uint32_t total, int threads,
                  float *combined, float *normal, float *albedo,
                  uint32_t *sample_count, oracle_stats *stats);

/* Same as oracle_render restricted to window rows [row_begin, row_end);
 * film pointers address the full window. */
int oracle_render_rows(const oracle_scene *scene,
                       const psx_render_settings *settings, uint32_t first,
                       uint32_t count, uint32_t total, int threads,
                       uint32_t row_begin, uint32_t row_end, float *combined,
                       float *normal, float *albedo, uint32_t *sample_count,
                       oracle_stats *stats);

/* Light-pass sums per pixel, in the reference's LightPassBuffer order
 * (src/luisa/path_tracer_internal.h:410-424): diffuse direct/indirect,
 * glossy direct/indirect, transmission direct/indirect, volume
 * direct/indirect, emission, environment, glossy colour, transmission
 * colour; each a float4 (rgb sum, sample count in w). */
#define ORACLE_LIGHT_PASSES 12u

/* oracle_render_rows that also accumulates the light passes into `light`
 * (ORACLE_LIGHT_PASSES float4 per window pixel; NULL skips them). */
int oracle_render_light(const oracle_scene *scene,
                        const psx_render_settings *settings, uint32_t first,
                        uint32_t count, uint32_t total, int threads,
                        uint32_t row_begin, uint32_t row_end, float *combined,
                        float *normal, float *albedo, uint32_t *sample_count,
                        float *light, oracle_stats *stats);

/* A light pass from the accumulated sums, as write_passes computes it
 * (src/luisa/path_tracer_session.cpp:150-230): emission, environment and
 * volume passes x exposure / N, colour passes / N, diffuse / glossy /
 * transmission passes divided by the albedo / glossy colour / transmission
 * colour sums channel by channel (safe_divide_even_color) x exposure. */
void oracle_write_light_pass(uint32_t pass_kind, uint32_t channels,
                             uint32_t pixel_count, float exposure,
                             const float *light, const float *albedo,
                             const uint32_t *sample_count, float *out);

/* Normalise accumulated film into a pass, as write_passes does
 * (src/luisa/path_tracer_session.cpp:71-254).  out holds `channels` floats
 * per pixel. */
void oracle_write_pass(uint32_t pass_kind, uint32_t channels,
                       uint32_t pixel_count, float exposure,
                       const float *combined, const float *normal,
                       const float *albedo, const uint32_t *sample_count,
                       float *out);

/* Primary probe: per window pixel (object, primitive, t bits, hit). */
int oracle_primary_hits(const oracle_scene *scene,
                        const psx_render_settings *settings,
                        uint32_t sample_index, uint32_t total,
                        uint32_t *out_records);

/* Path event log, PSX_TRACE_EVENTS x PSX_TRACE_WORDS u32 per pixel. */
int oracle_trace_events(const oracle_scene *scene,
                        const psx_render_settings *settings,
                        uint32_t sample_index, uint32_t total,
                        uint32_t *out_records);
/* schema v2 path trace of one pixel (window-relative) and sample:
 * PSX_PATH_TRACE_SLOTS float4 slots and, optionally, the pixel's events */
int oracle_path_trace(const oracle_scene *scene, const psx_render_settings *settings,
                      uint32_t x, uint32_t y, uint32_t sample_index, uint32_t total,
                      float *out_slots, uint32_t *out_records);

/* Sampler probe with the same layout as psx_session_sampler_probe. */
int oracle_sampler_probe(const psx_render_settings *settings,
                         uint32_t sample_index, uint32_t total,
                         uint32_t first_dimension, uint32_t dimension_count,
                         uint32_t *out);

/* Closest-hit and shadow queries with the layouts of psx_scene_intersect
 * (the oracle's shadow loop always walks every transparent hit in order). */
int oracle_scene_intersect(const oracle_scene *scene, const psx_ray_query *queries, uint32_t count,
                           psx_ray_hit *hits);

/* Single-function probes with the record layouts of psx_backend_probe. */
uint32_t oracle_probe_input_width(uint32_t function);
uint32_t oracle_probe_output_width(uint32_t function);
int oracle_probe(uint32_t function, const float *inputs, uint32_t count, float *outputs);

/* The closure fixture records of one diffuse closure (44 floats). */
int oracle_diffuse_closure_probe(const float color[3], float u_lobe, const float u_dir[2], float *out);

#ifdef __cplusplus
}
#endif

#endif
