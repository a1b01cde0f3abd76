/*
 * psx.h - C ABI of the MI355X render path for Psycles scenes.
 *
 * This is the drop-in boundary for Psycles' device render call.  Every entry
 * point replaces one member of the reference's C++ contract:
 *
 *   psx_backend_create / psx_backend_destroy
 *       luisa_backend::LuisaPathTracerBackend(Device, LuisaPathTracerOptions)
 *       (include/psycles/luisa/path_tracer.h:45-78)
 *   psx_scene_compile / psx_scene_destroy
 *       contract::RendererBackend::compile_scene(const SceneSnapshot&)
 *       -> SceneCompilation{unique_ptr<CompiledScene>, diagnostics}
 *       (include/psycles/contract/render.h:172-181, 145-161)
 *   psx_scene_update_parameters
 *       the parameter-only rebind branch of MaterialLibrary::update
 *       (src/compiler/material_library.cpp:53-90)
 *   psx_session_create / psx_session_destroy
 *       contract::RendererBackend::create_session(const CompiledScene&,
 *       const RenderSettings&) (include/psycles/contract/render.h:178-180;
 *       validation in src/luisa/path_tracer.cpp:22-75)
 *   psx_session_reset   RenderSession::reset (render.h:166)
 *   psx_session_render  RenderSession::render_samples(const SampleRange&,
 *                       OutputSink&) (render.h:167;
 *                       src/luisa/path_tracer_session.cpp:272-413)
 *   psx_session_cancel  RenderSession::cancel (render.h:168;
 *                       src/luisa/path_tracer_session.cpp:409-416)
 *
 * Conventions: opaque handles, int status codes (PSX_OK == 0, negative on
 * error), no exception ever crosses this boundary, no torch or HIP types in
 * any signature.  All scene arrays are plain host pointers that are only read
 * during the call that receives them.  Matrices are 4x4 column-major float
 * arrays exactly like psycles::Mat4f (elements[12..14] hold translation).
 */
#ifndef PSX_H
#define PSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSX_ABI_VERSION 2u

/* ---- status codes ------------------------------------------------------ */
enum {
    PSX_OK = 0,
    PSX_ERROR_INVALID_ARGUMENT = -1,
    PSX_ERROR_DEVICE = -2,        /* HIP runtime or kernel launch failure   */
    PSX_ERROR_OUT_OF_MEMORY = -3,
    PSX_ERROR_CANCELLED = -4,     /* render_samples returned false: cancel   */
    PSX_ERROR_RANGE = -5,         /* render_samples returned false: range    */
    PSX_ERROR_INCOMPLETE = -6,    /* sample_count readback mismatch         */
    PSX_ERROR_UNSUPPORTED = -7,
    PSX_ERROR_NO_DEVICE = -8
};

/* ---- enums mirroring the reference contract ----------------------------- */

/* contract::PassKind (include/psycles/contract/render.h:15-42).  The values
 * are identical so a caller can forward its own enum unchanged. */
enum psx_pass_kind {
    PSX_PASS_COMBINED = 0,
    PSX_PASS_DEPTH = 1,
    PSX_PASS_NORMAL = 2,
    PSX_PASS_POSITION = 3,
    PSX_PASS_ALBEDO = 4,
    PSX_PASS_EMISSION = 5,
    PSX_PASS_DIFFUSE_DIRECT = 6,
    PSX_PASS_DIFFUSE_INDIRECT = 7,
    PSX_PASS_GLOSSY_DIRECT = 8,
    PSX_PASS_GLOSSY_INDIRECT = 9,
    PSX_PASS_TRANSMISSION_DIRECT = 10,
    PSX_PASS_TRANSMISSION_INDIRECT = 11,
    PSX_PASS_VOLUME_DIRECT = 12,
    PSX_PASS_VOLUME_INDIRECT = 13,
    PSX_PASS_DENOISING_NORMAL = 18,
    PSX_PASS_DENOISING_ALBEDO = 19,
    PSX_PASS_SAMPLE_COUNT = 20,
    PSX_PASS_GLOSSY_COLOR = 23,
    PSX_PASS_TRANSMISSION_COLOR = 24,
    PSX_PASS_ENVIRONMENT = 25
};

/* contract::PixelFilter (render.h:65-69) */
enum psx_pixel_filter {
    PSX_FILTER_BOX = 0,
    PSX_FILTER_GAUSSIAN = 1,
    PSX_FILTER_BLACKMAN_HARRIS = 2
};

/* contract::DirectLightSampling (render.h:71-75) */
enum psx_direct_light_sampling {
    PSX_DIRECT_LIGHT_MIS = 0,
    PSX_DIRECT_LIGHT_FORWARD = 1,
    PSX_DIRECT_LIGHT_NEE = 2
};

/* contract::EmissionSampling (include/psycles/contract/scene.h:72-78) */
enum psx_emission_sampling {
    PSX_EMISSION_SAMPLING_NONE = 0,
    PSX_EMISSION_SAMPLING_AUTOMATIC = 1,
    PSX_EMISSION_SAMPLING_FRONT = 2,
    PSX_EMISSION_SAMPLING_BACK = 3,
    PSX_EMISSION_SAMPLING_FRONT_BACK = 4
};

/* contract::RayVisibility bits (scene.h:306-320) */
enum psx_visibility {
    PSX_VISIBILITY_CAMERA = 1u << 0,
    PSX_VISIBILITY_DIFFUSE = 1u << 1,
    PSX_VISIBILITY_GLOSSY = 1u << 2,
    PSX_VISIBILITY_TRANSMISSION = 1u << 3,
    PSX_VISIBILITY_SHADOW = 1u << 4,
    PSX_VISIBILITY_VOLUME_SCATTER = 1u << 5,
    PSX_VISIBILITY_ALL = 0x3fu
};

/* Closure-tree operations of a SurfaceProgram
 * (compiler::ClosureOperation, include/psycles/compiler/surface_program.h:383-396),
 * restricted to the closures this render path evaluates. */
enum psx_closure_op {
    PSX_CLOSURE_NULL = 0,
    PSX_CLOSURE_ADD = 1,
    PSX_CLOSURE_MIX = 2,
    PSX_CLOSURE_DIFFUSE = 3,
    PSX_CLOSURE_EMISSION = 4,
    PSX_CLOSURE_TRANSPARENT = 5
};

/* contract::LightType (scene.h:369-375) */
enum psx_light_type {
    PSX_LIGHT_POINT = 0,
    PSX_LIGHT_SPOT = 1,
    PSX_LIGHT_AREA = 2,
    PSX_LIGHT_DISTANT = 3,
    PSX_LIGHT_BACKGROUND = 4
};

/* ---- scene snapshot (flat form of contract::SceneSnapshot) -------------- */

/* One closure instruction.  Child and parameter fields are indices:
 *   a, b       child instruction indices (ADD, MIX)
 *   factor     scalar parameter index (MIX factor, clamped to [0,1])
 *   color      first of three consecutive parameter floats (DIFFUSE,
 *              EMISSION, TRANSPARENT)
 *   strength   scalar parameter index (EMISSION)
 *   roughness  scalar parameter index (DIFFUSE; 0 selects Lambert)
 * Unused fields must be zero.  Mirrors compiler::ClosureInstruction
 * (surface_program.h:398-440) with constant inputs lowered to parameters. */
typedef struct psx_closure_instr {
    uint32_t op;
    uint32_t a;
    uint32_t b;
    uint32_t factor;
    uint32_t color;
    uint32_t strength;
    uint32_t roughness;
    uint32_t reserved;
} psx_closure_instr;

/* A material = SurfaceProgram (instructions + root) + SurfaceParameterBlock.
 * Materials whose instruction arrays are equal share one compiled program
 * (the structure signature of docs/architecture.md:63-72); they differ only in
 * their parameter blocks. */
typedef struct psx_material {
    const psx_closure_instr *instructions;
    uint32_t instruction_count;
    uint32_t root;
    const float *parameters;
    uint32_t parameter_count;
    uint32_t emission_sampling; /* enum psx_emission_sampling */
} psx_material;

/* contract::TriangleMeshDesc (scene.h:182-239), triangle subset. */
typedef struct psx_mesh {
    const float *positions;          /* 3 * vertex_count                    */
    uint32_t vertex_count;
    const uint32_t *triangles;       /* 3 * triangle_count                  */
    uint32_t triangle_count;
    const float *normals;            /* optional point-domain normals, 3*V  */
    const uint8_t *triangle_smooth;  /* optional, one per triangle          */
    const uint32_t *material_slots;  /* material index per slot             */
    uint32_t material_slot_count;
    const uint32_t *triangle_material_slots; /* optional, one per triangle  */
} psx_mesh;

/* contract::InstanceDesc (scene.h:304-333). */
typedef struct psx_instance {
    uint32_t mesh;
    uint32_t visibility_mask;        /* enum psx_visibility bits            */
    float transform[16];
    const uint32_t *material_overrides;
    uint32_t material_override_count;
    /* Cycles' shadow-terminator geometry offset (>= 0, finite): shadow
     * origins on smooth triangles are lifted onto the parabolic surface
     * through the vertex normals (include/psycles/luisa/surface_ray.h:108-309). */
    float shadow_terminator_geometry_offset;
} psx_instance;

/* contract::LightDesc (scene.h:377-403): point (sphere or disk), spot,
 * distant, area (rectangle or ellipse, with spread) and background lights.
 * Background lights fold into the constant world radiance exactly like
 * src/luisa/path_tracer_scene.cpp:1728-1733; the others are sampled by NEE
 * through the flat CDF or the light tree and hit by forward rays. */
typedef struct psx_light {
    uint32_t type;
    uint32_t visibility_mask;
    float transform[16];
    float color[3];
    float power;
    float size;
    float size_y;
    float spread;
    float spot_angle;
    float spot_smooth;
    float angle;
    uint32_t flags;                  /* enum psx_light_flags               */
    uint32_t max_bounces;
} psx_light;

/* LightDesc booleans (scene.h:377-403). */
enum psx_light_flags {
    PSX_LIGHT_FLAG_NORMALIZE = 1u,
    PSX_LIGHT_FLAG_ELLIPSE = 2u,
    PSX_LIGHT_FLAG_SPHERE = 4u,
    PSX_LIGHT_FLAG_USE_MIS = 8u,
    PSX_LIGHT_FLAG_CAST_SHADOW = 16u
};

/* contract::CameraDesc (scene.h:347-367): perspective and orthographic
 * projections (CameraProjection, scene.h:335-339); no depth of field. */
typedef struct psx_camera {
    uint32_t projection;             /* 0 perspective, 1 orthographic       */
    uint32_t sensor_fit;             /* 0 horizontal, 1 vertical, 2 auto    */
    float transform[16];
    float field_of_view;
    float horizontal_field_of_view;
    float lens_shift_x;
    float lens_shift_y;
    float near_clip;
    float far_clip;
    float orthographic_scale;        /* view-plane extent of the fitted axis */
} psx_camera;

typedef struct psx_scene_desc {
    uint32_t abi_version;            /* PSX_ABI_VERSION                     */
    const psx_material *materials;
    uint32_t material_count;
    const psx_mesh *meshes;
    uint32_t mesh_count;
    const psx_instance *instances;
    uint32_t instance_count;
    const psx_light *lights;
    uint32_t light_count;
    psx_camera camera;
    /* Constant world shader radiance (Background color * strength); the
     * reference never puts a spatially constant world into the light
     * distribution (src/luisa/path_tracer_scene.cpp:1896-1911). */
    float world_color[3];
    uint32_t world_visibility_mask;
} psx_scene_desc;

/* ---- render settings (contract::RenderSettings, render.h:81-117) -------- */
typedef struct psx_integrator_settings {
    uint32_t max_bounces;
    uint32_t min_bounces;
    uint32_t diffuse_bounces;
    uint32_t glossy_bounces;
    uint32_t transmission_bounces;
    uint32_t volume_bounces;
    uint32_t transparent_min_bounces;
    uint32_t transparent_max_bounces;
    float sample_clamp_direct;
    float sample_clamp_indirect;
    float filter_glossy;
    float film_exposure;
    float light_sampling_threshold;
    uint32_t reflective_caustics;
    uint32_t refractive_caustics;
    uint32_t use_light_tree;
    uint32_t direct_light_sampling;  /* enum psx_direct_light_sampling      */
} psx_integrator_settings;

typedef struct psx_render_settings {
    uint32_t width;                  /* full extent                         */
    uint32_t height;
    uint32_t window_x;               /* pixel window; width 0 = full frame  */
    uint32_t window_y;
    uint32_t window_width;
    uint32_t window_height;
    uint32_t seed;
    uint32_t transparent_background;
    uint32_t pixel_filter;           /* enum psx_pixel_filter               */
    float filter_width;
    float pass_alpha_threshold;
    psx_integrator_settings integrator;
    const uint32_t *passes;          /* enum psx_pass_kind per request      */
    const uint32_t *pass_channels;   /* channels per request                */
    uint32_t pass_count;
} psx_render_settings;

/* Default values of contract::RenderSettings / PathIntegratorSettings. */
void psx_render_settings_default(psx_render_settings *settings);

/* ---- output sink (contract::OutputSink, render.h:136-143) --------------- */
typedef struct psx_sink {
    void *user;
    void (*begin)(void *user, const psx_render_settings *settings);
    /* pixels are row-major over the window, `channels` floats per pixel;
     * valid only for the duration of the call (PassTile::pixels). */
    void (*write)(void *user, uint32_t pass_kind, uint32_t channels,
                  uint32_t window_x, uint32_t window_y,
                  uint32_t window_width, uint32_t window_height,
                  uint32_t full_width, uint32_t full_height,
                  const float *pixels);
    void (*end)(void *user, int cancelled);
} psx_sink;

/* Diagnostic callback (contract::RenderDiagnostic). */
typedef void (*psx_diag_fn)(void *user, const char *message);

/* ---- handles ------------------------------------------------------------ */
typedef struct psx_backend psx_backend;
typedef struct psx_scene psx_scene;
typedef struct psx_session psx_session;

/* Backend options (luisa_backend::LuisaPathTracerOptions,
 * include/psycles/luisa/path_tracer.h:45-59). */
enum psx_integrator {
    PSX_INTEGRATOR_MEGAKERNEL = 0, /* one lane runs whole paths (plain path)  */
    PSX_INTEGRATOR_WAVEFRONT = 1   /* trace/shade kernels over HBM path state */
};

enum psx_backend_flags {
    PSX_BACKEND_COUNT_TRAVERSAL = 1u, /* count node visits and triangle tests */
    /* Launch every kernel of a wavefront dispatch from the host.  By default
     * a session captures its dispatch (counter reset, camera rays, every
     * bounce iteration's trace and shade, accumulation) into a graph once
     * and replays it per dispatch; the film is the same either way.  A
     * capture and the render server's device calls on the same backend never
     * overlap.  Device calls that other threads of the caller make during a
     * capture (allocations, synchronous copies) can invalidate it; callers
     * that drive sessions from several threads set this flag. */
    PSX_BACKEND_NO_GRAPHS = 2u,
    /* Time every wavefront trace launch with HIP events (psx_render_stats
     * trace_ms and trace_launches).  Events recorded inside a graph cannot be
     * timed on this runtime, so timed dispatches launch from the host and the
     * host waits for each dispatch to read its events. */
    PSX_BACKEND_TIME_TRACE = 4u,
    /* Shade wavefront dispatches with generated surface programs: one
     * straight-line closure program per distinct closure-leaf layout of the
     * scene's materials, compiled with hipRTC on the first dispatch that
     * needs it (one module per layout set and shade variant, cached by the
     * backend; parameter-only rebinds reuse it).  Films are bit-identical to
     * the built-in interpreter's. */
    PSX_BACKEND_GENERATED_PROGRAMS = 8u,
    /* Branching factor of the BVHs the backend's scene compiles build: two
     * or three children per node instead of four (the binary SAH tree is
     * collapsed to that width; records stay 64-byte 4-wide nodes with unused
     * slots empty).  Setting both is PSX_ERROR_INVALID_ARGUMENT.  Hits,
     * films and events are the same at every width (closest hits resolve
     * ties by primitive order, not by the tree); only traversal time
     * changes.  psycles_amd.tuning.tune_bvh_width picks the width by timing
     * probe dispatches. */
    PSX_BACKEND_BVH_WIDTH_2 = 16u,
    PSX_BACKEND_BVH_WIDTH_3 = 32u
};

typedef struct psx_backend_options {
    uint32_t next_event_estimation;  /* default 1                           */
    uint32_t max_samples_per_dispatch; /* default 4; batches kept in order  */
    uint32_t integrator;             /* enum psx_integrator                 */
    uint32_t flags;                  /* enum psx_backend_flags              */
} psx_backend_options;

const char *psx_version(void);
int psx_device_count(void);
/* The most recent HIP failure behind a PSX_ERROR_DEVICE or
 * PSX_ERROR_OUT_OF_MEMORY on the calling thread ("<HIP error name> at
 * <source>:<line>"), or "" if there was none.  Valid until the thread's next
 * call into the library. */
const char *psx_last_error(void);

int psx_backend_create(int device_ordinal, const psx_backend_options *options,
                       psx_backend **out_backend);
void psx_backend_destroy(psx_backend *backend);

/* ---- generated surface programs (PSX_BACKEND_GENERATED_PROGRAMS) ---------
 * The reference JIT-compiles each SurfaceProgram per structure signature
 * and keeps socket values in parameter buffers (docs/architecture.md:63-110,
 * include/psycles/compiler/surface_program.h:512-568).  These entry points
 * expose the generator without a device: the HIP source a set of materials
 * generates for a shade variant, and a hipRTC build of it. */
enum psx_program_variant {
    PSX_PROGRAM_LIGHT_PASSES = 1u, /* shade kernel that sums light passes */
    PSX_PROGRAM_LIGHTS = 2u        /* shade kernel for a scene with lights */
};
/* Writes the NUL-terminated source into `source` (may be NULL), sets
 * `length` (without the NUL) and `programs` (distinct closure-leaf layouts;
 * may be NULL).  PSX_ERROR_INVALID_ARGUMENT for an invalid material, a
 * variant above 3, or a capacity <= length with a non-NULL `source`. */
int psx_surface_program_source(const psx_material *materials, uint32_t material_count, uint32_t variant,
                               char *source, size_t capacity, size_t *length, uint32_t *programs);
/* Compiles that source with hipRTC for `arch` (NULL: "gfx950").  PSX_OK and
 * the code-object size on success; PSX_ERROR_UNSUPPORTED when it does not
 * compile.  The compiler log goes to `log` (may be NULL). */
int psx_surface_program_build(const psx_material *materials, uint32_t material_count, uint32_t variant,
                              const char *arch, char *log, size_t log_capacity, uint64_t *code_bytes);
/* Generated modules the backend has compiled so far. */
uint32_t psx_backend_generated_programs(const psx_backend *backend);

int psx_scene_compile(psx_backend *backend, const psx_scene_desc *desc,
                      psx_diag_fn diagnostics, void *diagnostics_user,
                      psx_scene **out_scene);
void psx_scene_destroy(psx_scene *scene);
/* Replace the parameter block of one material without recompiling its
 * program.  Fails with PSX_ERROR_INVALID_ARGUMENT when parameter_count
 * differs from the compiled block (a structure change needs a recompile). */
int psx_scene_update_parameters(psx_scene *scene, uint32_t material,
                                const float *parameters,
                                uint32_t parameter_count);

int psx_session_create(psx_scene *scene, const psx_render_settings *settings,
                       psx_session **out_session);
void psx_session_destroy(psx_session *session);
int psx_session_reset(psx_session *session,
                      const psx_render_settings *settings);
/* SampleRange{first, count, offset, total}.  Returns PSX_OK when the sink
 * received every requested pass, otherwise the reason render_samples would
 * have returned false. */
int psx_session_render(psx_session *session, uint32_t first, uint32_t count,
                       uint32_t offset, uint32_t total, const psx_sink *sink);
/* Thread-safe. */
void psx_session_cancel(psx_session *session);

/* ---- measurement and inspection ---------------------------------------- */

typedef struct psx_render_stats {
    uint64_t rays;            /* camera + continuation + shadow re-queries  */
    uint64_t samples;         /* completed pixel-samples                    */
    uint64_t dispatches;      /* kernel launches of the render loop         */
    double kernel_ms;         /* device time of the render kernels          */
    uint64_t node_visits;     /* BVH node records fetched (0 if disabled)   */
    uint64_t triangle_tests;  /* triangle records tested (0 if disabled)    */
    double trace_ms;          /* wavefront with PSX_BACKEND_TIME_TRACE: device
                                 time of the trace kernel (else 0)          */
    uint64_t trace_launches;  /* ... and its launches (else 0)              */
    uint64_t graph_launches;  /* wavefront: dispatches replayed as a graph  */
    uint64_t graph_builds;    /* wavefront: graphs captured by this call    */
} psx_render_stats;

/* Statistics of the most recent psx_session_render. */
int psx_session_stats(const psx_session *session, psx_render_stats *stats);

/* Runs the render loop on samples [first, first+count) without reading the
 * film back or calling a sink: the timed region of the benchmark.  The film
 * keeps accumulating exactly like psx_session_render. */
int psx_session_render_device(psx_session *session, uint32_t first,
                              uint32_t count, uint32_t total);

/* ---- wavefront launch shape and its autotuner ----------------------------
 * The wavefront integrator's launch shape.  Every value in range gives the
 * same film, bit for bit; only the speed changes.  The closest the reference
 * has is its fixed dispatch policy (backend_max_pixel_samples_per_dispatch,
 * src/luisa/path_tracer_backend_policy.h:13-22); these are the MI355X
 * kernels' own knobs. */
typedef struct psx_wavefront_tuning {
    uint32_t tile_width;   /* 8, 16, 32 or 64: a wave's 64 camera paths cover
                              tile_width x (64 / tile_width) pixels (8)      */
    uint32_t refill_lanes; /* 1..64: idle lanes of a trace wave that make it
                              take new rays from the queue (16)              */
    uint32_t trace_chunk;  /* 16..4096: queue entries a trace wave reserves
                              per atomic (128)                               */
    uint32_t queue_chunk;  /* 128 or 256: next-queue slots a shade wave
                              reserves per atomic (256)                      */
    uint32_t list_chunk;   /* 64, 128 or 256: shade-list slots a shade wave
                              reserves per atomic (128)                      */
    uint32_t sort_shading; /* 0 or 1: each shade block takes its 256 paths
                              sorted by closure structure of the hit (0)    */
    uint32_t xcd_segments; /* 1 or 8: trace waves take rays from one queue, or
                              from 8 segments, one per XCD, stealing from the
                              others when theirs is drained (1)             */
} psx_wavefront_tuning;

void psx_wavefront_tuning_default(psx_wavefront_tuning *tuning);
/* PSX_ERROR_INVALID_ARGUMENT for a value out of range. */
int psx_session_set_tuning(psx_session *session, const psx_wavefront_tuning *tuning);
int psx_session_get_tuning(const psx_session *session, psx_wavefront_tuning *tuning);

typedef struct psx_tuning_report {
    psx_wavefront_tuning chosen;
    float default_ms;        /* probe dispatch time with the default shape  */
    float chosen_ms;         /* probe dispatch time with the chosen shape   */
    uint32_t candidates;     /* shapes timed                                */
    uint32_t probe_samples;  /* samples per probe dispatch                  */
} psx_tuning_report;

/* Times probe dispatches of `probe_samples` samples (1..64) of the session's
 * window for candidate shapes, one knob at a time from the default (tile
 * width, refill lanes, trace chunk, shade chunks, shade sorting, XCD
 * segments), keeps a candidate only if
 * it is at least 1 % faster, then times the winner and the default shape
 * three more times in turns and keeps the winner only if it stays at least
 * 1 % faster; the result is set on the session.  The probes
 * render into scratch memory: the session's film, sample count and stats are
 * untouched.  A session whose renders run on the megakernel (integrator
 * PSX_INTEGRATOR_MEGAKERNEL without light passes) returns
 * PSX_ERROR_UNSUPPORTED. */
int psx_session_autotune(psx_session *session, uint32_t probe_samples,
                         psx_tuning_report *report);

/* Raw accumulated film, row-major over the window: combined float4 (rgb sum,
 * transparency sum), normal float4, albedo float4 and the u32 sample count.
 * Any pointer may be NULL. */
int psx_session_read_film(psx_session *session, float *combined,
                          float *normal, float *albedo,
                          uint32_t *sample_count);

/* Primary-ray probe: for every window pixel at sample `sample_index`, the
 * closest camera-ray hit as (object, primitive, t bits, hit flag). */
int psx_session_primary_hits(psx_session *session, uint32_t sample_index,
                             uint32_t total, uint32_t *out_records);

/* Path event log for one sample index: per pixel, PSX_TRACE_EVENTS records of
 * PSX_TRACE_WORDS u32 each (see psx_trace_word). */
#define PSX_TRACE_EVENTS 8u
#define PSX_TRACE_WORDS 8u
enum psx_trace_word {
    PSX_TRACE_FLAGS = 0,      /* bit0 valid, bit1 surface, bit2 background,
                                 bit3 rr_terminated, bit4 nee_attempted,
                                 bit5 scatter_valid, bit6 transparent      */
    PSX_TRACE_RNG_OFFSET = 1,
    PSX_TRACE_OBJECT = 2,
    PSX_TRACE_PRIMITIVE = 3,
    PSX_TRACE_LIGHT = 4,      /* selected distribution entry or ~0u         */
    PSX_TRACE_CLOSURE = 5,    /* selected closure index or ~0u              */
    PSX_TRACE_TERMINATE_BITS = 6, /* float bits of u_term                   */
    PSX_TRACE_RR_BITS = 7     /* float bits of continuation probability    */
};
int psx_session_trace_events(psx_session *session, uint32_t sample_index,
                             uint32_t total, uint32_t *out_records);

/* Diagnostic path trace of one pixel and sample: PSX_PATH_TRACE_SLOTS float4
 * slots in the "psycles.cycles-path-trace" schema version 2 layout
 * (include/psycles/luisa/path_trace_schema.h:10-121, written at
 * src/luisa/path_kernel_setup.cpp:540-560 and
 * src/luisa/path_kernel_surface_shading.cpp:270-340); replaces
 * LuisaPathTraceRequest{pixel_x, pixel_y, sample, sink}
 * (include/psycles/luisa/path_tracer.h:20-43).  `pixel_x` is a column of the
 * full image, `pixel_y` a row counted from the lower row of the full image
 * (Cycles' film convention); the pixel must lie in the session's window.
 * w = 1 marks a written slot.  Written: the global slots 0-6 and, for the
 * first 4 surface events, event slots 0-16, 32-36, 38 and 39.  `out_records`
 * (optional) receives the pixel's PSX_TRACE_EVENTS x PSX_TRACE_WORDS
 * discrete events. */
#define PSX_PATH_TRACE_SLOTS 328u
#define PSX_PATH_TRACE_VERSION 2u
int psx_session_path_trace(psx_session *session, uint32_t pixel_x, uint32_t pixel_y,
                           uint32_t sample_index, uint32_t total,
                           float *out_slots, uint32_t *out_records);

/* Sobol LUT as built for `total` samples (256 patterns x sequence size
 * float4); out may be NULL to query *out_count. */
int psx_sobol_table(uint32_t total, float *out, uint32_t *out_count);

/* Raw light-pass sums of a session whose settings request any light pass:
 * per window pixel, 12 float4 in the order diffuse direct/indirect, glossy
 * direct/indirect, transmission direct/indirect, volume direct/indirect,
 * emission, environment, glossy colour, transmission colour (rgb sum,
 * samples in w).  PSX_ERROR_UNSUPPORTED when no light pass was requested. */
int psx_session_read_light_passes(psx_session *session, float *out);

/* Per-pixel sampler probe on the device: for each pixel of the window and the
 * given sample index, the u32 bit patterns of sample_4d at `dimension_count`
 * consecutive dimensions starting at `first_dimension`, plus rng_hash. */
int psx_session_sampler_probe(psx_session *session, uint32_t sample_index,
                              uint32_t total, uint32_t first_dimension,
                              uint32_t dimension_count, uint32_t *out);

/* ---- ray queries ----------------------------------------------------------
 *
 * Closest-hit and shadow queries against a compiled scene: the traversal
 * contract SceneTraversalComponent::closest / closest_shadow
 * (src/luisa/path_kernel_scene_traversal.cpp:34-68, 212-526: closed
 * interval [tmin, tmax], source and light identities excluded, visibility
 * mask, equal distances resolved to the greatest (object, primitive)), and
 * for shadow queries the transparent shadow loop trace_shadow
 * (src/luisa/path_tracer_geometry.cpp:75-174).  Objects are instance
 * indices; primitives are scene primitive ids: the sum of the triangle
 * counts of the earlier meshes plus the triangle's index in its mesh. */
#define PSX_INVALID_ID 0xffffffffu
enum psx_query_kind {
    PSX_QUERY_CLOSEST = 0,  /* closest hit                                  */
    PSX_QUERY_SHADOW = 1    /* transmittance through transparent surfaces   */
};
typedef struct psx_ray_query {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
    uint32_t visibility;        /* psx_visibility bits an instance must share */
    uint32_t source_object;     /* excluded identity or PSX_INVALID_ID        */
    uint32_t source_primitive;
    uint32_t light_object;      /* second excluded identity                   */
    uint32_t light_primitive;
    uint32_t kind;              /* enum psx_query_kind                        */
    uint32_t transparent_depth; /* shadow: transparent bounces already taken  */
    uint32_t transparent_max;   /* shadow: transparent bounce limit           */
} psx_ray_query;                /* 64 bytes */

typedef struct psx_ray_hit {
    float t, u, v;              /* closest: distance and barycentrics         */
    uint32_t object;            /* closest: instance or PSX_INVALID_ID        */
    uint32_t primitive;         /* closest: scene primitive id                */
    uint32_t hit;               /* closest: 1 hit; shadow: 1 fully occluded   */
    uint32_t traversals;        /* traversal passes the query used            */
    uint32_t reserved;
    float transmittance[3];     /* shadow: product of transparent extinctions */
    uint32_t reserved2;
} psx_ray_hit;                  /* 48 bytes */

/* Host arrays in and out; synchronous.  A shadow query returns as soon as
 * any opaque surface lies in its interval (its transmittance is then exactly
 * zero); transparent surfaces before it are taken in distance order. */
int psx_scene_intersect(psx_scene *scene, const psx_ray_query *queries, uint32_t count, psx_ray_hit *hits);
/* The same with device-resident queries and hits (device pointers of the
 * scene's device); allocates its traversal scratch for the call and returns
 * when the hits are written. */
int psx_scene_intersect_device(psx_scene *scene, const psx_ray_query *device_queries, uint32_t count,
                               psx_ray_hit *device_hits);

/* psx_scene_intersect_device without any allocation.  No counterpart in
 * the reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Its traversal
 * scratch (a work cursor and a stack spill column per lane, for at most
 * 65,536 lanes) is one static device array of 16 MiB + 256 B per device,
 * resolved when a backend is created on that device; nothing is allocated per
 * call.  Every call on a device shares that array: calls that overlap on one
 * device race.  More than PSX_INTERSECT_NOALLOC_MAX_QUERIES queries is
 * PSX_ERROR_INVALID_ARGUMENT. */
#define PSX_INTERSECT_NOALLOC_MAX_QUERIES (1u << 30)
int psx_scene_intersect_device_noalloc(psx_scene *scene, const psx_ray_query *device_queries, uint32_t count,
                                       psx_ray_hit *device_hits);

/* ---- single-function probes ----------------------------------------------
 *
 * Evaluates one device function of the render path on `count` input records
 * (host arrays, synchronous), so that parity tests can hold the device code
 * itself against the reference's function fixtures.  Record layouts, all
 * float (flags 0 or 1, integers as exact floats):
 *   FILM_CLAMP          in  contribution[3], depth, direct limit,
 *                           indirect limit (6)
 *                       out contribution[3] (3)
 *                       (cycles_film_light::clamp,
 *                        include/psycles/luisa/cycles_film_light.h:14-50)
 *   T_OFFSET            in  distance (1) out next tmin (1)
 *                       (surface_ray::intersection_t_offset)
 *   SELF_EXCLUSION      in  world origin[3], world geometric normal[3],
 *                           intersection-space origin[3] and direction[3],
 *                           p0[3], p1[3], p2[3] (21)
 *                       out origin[3], certified (4)
 *                       (origin_with_explicit_self_exclusion)
 *   SHADOW_TERMINATOR   in  position[3], shading normal[3], geometric
 *                           normal[3], light direction[3], geometry offset,
 *                           smooth, transform applied, barycentric[2],
 *                           p0..p2[9], n0..n2[9], object-to-world affine[12]
 *                           (columns x, y, z, translation) (47)
 *                       out position[3], skip_self (4)
 *                       (shadow_terminator_origin,
 *                        include/psycles/luisa/surface_ray.h:183-243)
 *   SURFACE_SHADOW_ORIGIN
 *                       in  the SHADOW_TERMINATOR record, then the
 *                           world-to-object affine[12] (59)
 *                       out position[3], skip_self (4)
 *                       (surface_shadow_origin: the lift, then the source
 *                        triangle certificate while skip_self,
 *                        include/psycles/luisa/surface_ray.h:252-309)
 *   TRIANGLE_SAMPLE     in  reference point[3], p0..p2[9], random[2] (14)
 *                       out position[3], direction[3], distance, pdf,
 *                           valid, solid angle used (10)
 *                       (triangle_light_sample<false>,
 *                        include/psycles/luisa/triangle_light_sampling.h)
 *   TRIANGLE_PDF        in  reference point[3], light point[3], p0..p2[9] (15)
 *                       out pdf, valid, solid angle used (3)
 *   DIFFUSE_EVAL        in  normal[3], roughness, colour[3], incoming[3],
 *                           outgoing[3] (13)
 *                       out cosine-weighted intensity[3], pdf (4)
 *                       (Lambert / Oren-Nayar,
 *                        src/luisa/graph_surface_scattering.cpp:463-506)
 *   AREA_SAMPLE         in  lamp centre[3], unit axes x[3], y[3], z[3],
 *                           edge lengths u and v, spread, shape (bit 0
 *                           ellipse, bit 1 normalized power), reference
 *                           point[3], random[2] (21)
 *                       out position[3], direction[3], distance, pdf,
 *                           evaluation factor, valid (10)
 *                       (AreaLightSampling::from_position,
 *                        src/luisa/area_light_sampling.cpp)
 *   VOLUME              homogeneous volume transport (32 in, 24 out; unused
 *                       floats are zero).  in[0] selects the function; a
 *                       coefficient block C is sigma_t[3], sigma_s[3],
 *                       emission[3], flags (bit 0 extinction, bit 1
 *                       scatter, bit 2 emission):
 *                       0 transmittance: sigma_t[3], distance -> [3]
 *                       1 emission integral: C, distance -> [3]
 *                       2 channel pdf: albedo[3], throughput[3] -> [3]
 *                       3 channel sample: albedo[3], throughput[3], random
 *                         -> channel, rescaled random, matched, pdf[3]
 *                       4 bounded exponential sample: random, rate, min,
 *                         max -> distance
 *                       5 bounded exponential pdf: distance, rate[3], min,
 *                         max -> [3]
 *                       6 sample: C, distance, throughput[3], scatter
 *                         random, channel random, terminate;
 *                       7 the same with scatter probability[3] appended
 *                         -> transmittance[3], emission[3], throughput[3],
 *                         scatter probability[3], channel pdf[3],
 *                         distance, event pdf, scatter random, reservoir
 *                         random, channel, scattered, active
 *                       8 scatter probability: C, distance, terminate,
 *                         scattered radiance[3], transmitted radiance[3],
 *                         majorant optical depth, guiding enabled -> [3]
 *                       9 direct sampling preparation: method, random,
 *                         enabled -> method, random, use MIS, enabled
 *                       10 equiangular sample: origin[3], direction[3],
 *                         light position[3], min, max, random -> distance,
 *                         pdf, valid
 *                       11 equiangular pdf: as 10 with distance for random
 *                         -> pdf
 *                       12 power heuristic: pdf, other pdf -> weight
 *                       13 direct-light scatter point on a ray from the
 *                         origin along +z: C, distance, throughput[3],
 *                         scatter random, reservoir random, requested
 *                         method (0 none, 1 distance, 2 equiangular, 3
 *                         MIS), technique random, light position[3],
 *                         min, max -> transmittance[3], throughput[3],
 *                         channel pdf[3], distance, distance pdf,
 *                         equiangular pdf, MIS weight, channel, method,
 *                         use MIS, scattered
 *                       14 spot light interval: ray origin[3], direction[3],
 *                         min, max, lamp centre[3], axes x[3], y[3], z[3],
 *                         axis scale[3], radius, spot angle -> min, max,
 *                         valid
 *                       15 area light interval: ray origin[3], direction[3],
 *                         min, max, centre[3], axis u[3], v[3], z[3],
 *                         length u, length v, spread, ellipse -> min,
 *                         max, valid
 *                       16 spot lamp sample for a volume segment and at a
 *                         collision: reference point[3], centre[3],
 *                         radius, sphere, axes x[3], y[3], z[3], axis
 *                         scale[3], random[2], normalized power, spot
 *                         angle, spot smoothness -> segment valid,
 *                         evaluation, pdf, collision valid, evaluation,
 *                         pdf
 *                       17 Henyey-Greenstein phase: anisotropy, cosine,
 *                         axis[3], random[2] -> pdf at the cosine, clamped
 *                         anisotropy, sampled direction[3], its pdf,
 *                         Rayleigh pdf at the cosine, Rayleigh sampled
 *                         direction[3] and its pdf
 *                       (src/luisa/homogeneous_volume_transport.cpp,
 *                        src/luisa/volume_scatter_probability.cpp:20-98,
 *                        src/luisa/volume_direct_sampling.cpp,
 *                        src/luisa/volume_light_interval.cpp,
 *                        src/luisa/volume_analytic_light_sampling.cpp,
 *                        include/psycles/luisa/cycles_volume_phase.h) */
enum psx_probe_function {
    PSX_PROBE_FILM_CLAMP = 0,
    PSX_PROBE_T_OFFSET = 1,
    PSX_PROBE_SELF_EXCLUSION = 2,
    PSX_PROBE_SHADOW_TERMINATOR = 3,
    PSX_PROBE_TRIANGLE_SAMPLE = 4,
    PSX_PROBE_TRIANGLE_PDF = 5,
    PSX_PROBE_DIFFUSE_EVAL = 6,
    PSX_PROBE_SURFACE_SHADOW_ORIGIN = 7,
    PSX_PROBE_AREA_SAMPLE = 8,
    PSX_PROBE_VOLUME = 9
};
/* Input and output floats per record of a probe function; 0 if unknown. */
uint32_t psx_probe_input_width(uint32_t function);
uint32_t psx_probe_output_width(uint32_t function);
int psx_backend_probe(psx_backend *backend, uint32_t function, const float *inputs, uint32_t count,
                      float *outputs);

/* ---- render server ------------------------------------------------------
 *
 * A request loop over resident compiled scenes (the serving form of
 * RenderSession::render_samples; the reference has no server, so these entry
 * points have no single counterpart: each request is a SampleRange of one
 * RenderSettings on one CompiledScene, with render_samples' result contract).
 *
 * Requests are taken in submission order by one worker thread.  A request
 * binds the scene's material parameters as they are when the worker takes it,
 * so psx_server_update_parameters affects exactly the requests submitted
 * after it.  Admitted requests (film allocated within the HBM budget) run as
 * chunks of their sample range inside batches: one wavefront launch sequence
 * over the paths of every request in the batch that shares its scene,
 * parameter version and integrator settings.  Batches run on several streams
 * at once; viewport requests (window pixels and sample count within the
 * options' limits) are batched ahead of the others and have a stream of their
 * own; while viewport work waits or runs no final-frame batch starts and
 * running ones launch their kernels on a quarter of the compute units.
 * While viewport requests were seen within the last second, final-frame
 * batches are sized from the measured device time per path of their scene
 * and settings to last about target_batch_ms; otherwise they fill the
 * path capacity.  A request's film is accumulated in sample order, one chunk
 * at a time, so it is bit-identical to a session rendering the same range
 * alone.
 * All entry points are thread-safe. */

typedef struct psx_server psx_server;

typedef struct psx_server_options {
    uint32_t path_capacity;        /* paths per batch (0: 8,388,608)        */
    uint32_t streams;              /* concurrent batches, 2..8 (0: 6); the  */
                                   /* upper half take viewport batches only */
    uint64_t hbm_budget;           /* bytes for films + batch state (0: 200 GB) */
    uint32_t viewport_max_pixels;  /* (0: 256,000)                          */
    uint32_t viewport_max_samples; /* (0: 4)                                */
    uint32_t target_batch_ms;      /* final-frame batch length the server   */
                                   /* sizes batches for while viewport work */
                                   /* is around (0: 12)                     */
    uint32_t reserved[3];
} psx_server_options;

enum psx_request_state {
    PSX_REQUEST_QUEUED = 0,    /* waiting for HBM admission                 */
    PSX_REQUEST_RUNNING = 1,   /* admitted; chunks in progress              */
    PSX_REQUEST_DONE = 2,
    PSX_REQUEST_CANCELLED = 3,
    PSX_REQUEST_FAILED = 4
};

typedef struct psx_request_status {
    uint32_t state;            /* psx_request_state                         */
    uint32_t samples_done;     /* samples accumulated so far                */
    uint32_t batches;          /* batches the request took part in          */
    uint32_t viewport;         /* 1 when batched as a viewport request      */
    double submit_ms;          /* server clock (ms since server creation)   */
    double admit_ms;
    double first_launch_ms;
    double done_ms;            /* latency = done_ms - submit_ms             */
} psx_request_status;

typedef struct psx_serving_stats {
    uint64_t submitted, completed, batches, paths, rays;
    uint64_t edits;            /* parameter versions created               */
    uint64_t hbm_in_use;       /* films + batch state, bytes                */
    uint64_t peak_hbm;
    double busy_ms;            /* wall time with at least one batch running */
} psx_serving_stats;

int psx_server_create(psx_backend *backend, const psx_server_options *options,
                      psx_server **out);
/* Cancels queued work, waits for running batches, frees every request.
 * Every scene a request or edit named must still exist: destroy the server
 * before those scenes. */
void psx_server_destroy(psx_server *server);
/* Queues samples [first, first+count) of `total` (SampleRange, offset 0) of
 * `settings` on `scene`.  Validation as psx_session_create and
 * render_samples: PSX_ERROR_INVALID_ARGUMENT / PSX_ERROR_RANGE; settings
 * that request light passes are PSX_ERROR_UNSUPPORTED (sessions render
 * them). */
int psx_server_submit(psx_server *server, psx_scene *scene,
                      const psx_render_settings *settings, uint32_t first,
                      uint32_t count, uint32_t total, uint64_t *out_id);
/* Parameter-only material edit ordered with submissions (see above). */
int psx_server_update_parameters(psx_server *server, psx_scene *scene,
                                 uint32_t material, const float *parameters,
                                 uint32_t parameter_count);
int psx_server_status(psx_server *server, uint64_t id, psx_request_status *out);
/* Blocks until the request is done, cancelled or failed, or timeout_ms
 * passes (negative: no limit); returns PSX_OK when done.  A request whose
 * film cannot fit the HBM budget even with no other film resident fails
 * with PSX_ERROR_OUT_OF_MEMORY; a device error on the server's threads fails
 * every request with its status.  psx_last_error then names the cause. */
int psx_server_wait(psx_server *server, uint64_t id, double timeout_ms);
/* A done request's film: raw sums as psx_session_read_film, or passes
 * through a sink as psx_session_render (begin, one tile per pass, end). */
int psx_server_read_film(psx_server *server, uint64_t id, float *combined,
                         float *normal, float *albedo, uint32_t *counts);
int psx_server_write_passes(psx_server *server, uint64_t id,
                            const psx_sink *sink);
int psx_server_cancel(psx_server *server, uint64_t id);
/* Frees a finished request and its film. */
int psx_server_release(psx_server *server, uint64_t id);
int psx_server_stats(psx_server *server, psx_serving_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* PSX_H */
