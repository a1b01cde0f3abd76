// psx_device.h - device record layouts of the MI355X render path.
//
// Shared by the host scene compiler (psx_host.cpp) and the HIP kernels
// (psx_path.h, psx_kernels.hip, psx_wavefront.hip).  Every record is sized for coalesced HBM fetches:
// BVH4 nodes are one 128-byte line, leaf triangles are three float4 (48 B).
#pragma once

#ifdef __HIPCC_RTC__
// Run-time compilation (hipRTC) of generated surface programs: hipRTC
// provides the HIP runtime and vector types but no C library headers.
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#define FLT_MAX __FLT_MAX__
#define FLT_MIN __FLT_MIN__
#define FLT_EPSILON __FLT_EPSILON__
#define INFINITY __builtin_huge_valf()
#else
#include <hip/hip_vector_types.h>

#include <cstdint>
#endif

namespace psx {

// ---------------------------------------------------------------------------
// acceleration structure
// ---------------------------------------------------------------------------

// 4-wide BVH node, one 128-byte line.  Child boxes are stored structure-of-
// arrays so one node fetch feeds four slab tests.  count[i]:
//   0            inner child, child[i] = node index
//   1..          leaf: BLAS nodes hold count[i] triangles starting at
//                child[i]; TLAS nodes hold one instance, child[i] = instance
//   kEmptySlot   unused slot
struct alignas(16) Node4 {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t count[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");
constexpr uint32_t kEmptySlot = 0xffffffffu;

// Traversal entries (stack entries and quantised child references) carry
// their kind in the two top bits.
constexpr uint32_t kKindMask = 0xc0000000u;
constexpr uint32_t kKindBlas = 0x00000000u;      // BLAS inner node index
constexpr uint32_t kKindTlas = 0x40000000u;      // TLAS inner node index
constexpr uint32_t kKindLeaf = 0x80000000u;      // first triangle << 3 | (count - 1)
constexpr uint32_t kKindInstance = 0xc0000000u;  // instance index

// The node record the kernels traverse: a Node4 quantised to 64 bytes, so a
// node fetch moves half the bytes and the L2 holds twice the nodes.  Child
// bound = origin + q * 2^exp per axis, with q an 8-bit integer rounded
// outward (bounds only grow, so culling stays conservative).  ref[i] is the
// child's traversal entry (kind | index), kEmptySlot for an unused slot.
struct alignas(16) QNode4 {
    float origin[3];
    int8_t exp[3];
    uint8_t pad0;
    uint8_t lo_x[4], hi_x[4], lo_y[4], hi_y[4];
    uint8_t lo_z[4], hi_z[4];
    uint32_t pad1[2];
    uint32_t ref[4];
};
static_assert(sizeof(QNode4) == 64, "QNode4 must be 64 bytes");

// Leaf triangle in intersection space (object space for shared meshes, final
// world space for single-user meshes).  p0.w carries the local primitive
// index as raw bits.
struct alignas(16) LeafTri {
    float p0x, p0y, p0z, prim_bits;
    float p1x, p1y, p1z, pad1;
    float p2x, p2y, p2z, pad2;
};
static_assert(sizeof(LeafTri) == 48, "LeafTri must be 48 bytes");

// Affine transform stored column-major without the constant last row:
// m[0..2] column 0, m[3..5] column 1, m[6..8] column 2, m[9..11] translation.
struct Affine {
    float m[12];
};

struct alignas(16) Instance {
    Affine w2o;            // cycles_inverse_transform(transform)
    Affine o2w;            // transform
    Affine n2w;            // transpose(w2o) (translation column unused)
    uint32_t blas_root;    // node index of the BLAS root in the node array
    uint32_t mesh;
    uint32_t visibility;
    uint32_t applied;      // 1: positions are final world support positions
    uint32_t prim_offset;  // Cycles primitive offset of the mesh
    uint32_t ipos_offset;  // first intersection-space vertex (float3 index)
    uint32_t override_offset;
    uint32_t override_count;
    float terminator_offset;  // shadow_terminator_geometry_offset
    uint32_t shade_class;     // sort key of the materials it can use, 1..kShadeKeys-1 (0: miss)
    uint32_t pad[2];
};

// Sort keys of shading work: 0 for a miss, then one per closure structure
// (materials whose closure trees have the same shape run the same code), the
// last one for instances whose materials differ in structure and for
// structures beyond the first kShadeKeys - 2.
constexpr uint32_t kShadeKeys = 8u;

struct Mesh {
    uint32_t vertex_offset;    // object-space positions (float3 index)
    uint32_t index_offset;     // first triangle (uint3 index)
    uint32_t normal_offset;    // float3 index or kNone
    uint32_t smooth_offset;    // byte index or kNone
    uint32_t slot_offset;      // material slot table
    uint32_t slot_count;
    uint32_t tri_slot_offset;  // per-triangle slot or kNone
    uint32_t triangle_count;
};
constexpr uint32_t kNone = 0xffffffffu;

// entries of the per-lane traversal stack (LDS short stack plus HBM spill);
// scene compilation rejects trees whose worst-case stack use exceeds it.
constexpr uint32_t kTraversalStackDepth = 64u;
// leaf stack entries pack (first triangle << 3 | count - 1) under two kind bits
constexpr uint32_t kMaxLeafTriangles = 1u << 27;

// ---------------------------------------------------------------------------
// materials: a SurfaceProgram lowered to a flat leaf list.  A leaf's mix
// weight is the product, in root-to-leaf order, of one factor per enclosing
// mix: (1 - f) for the first operand, f for the second.
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxLeafDepth = 6;
constexpr uint32_t kMaxLeaves = 8;

enum LeafOp : uint32_t { kLeafDiffuse = 3u, kLeafEmission = 4u, kLeafTransparent = 5u };

struct Leaf {
    uint32_t op;
    uint32_t color;           // parameter index of the rgb triple
    uint32_t aux;             // strength (emission) or roughness (diffuse)
    uint32_t depth;           // number of enclosing mixes
    uint32_t factor[kMaxLeafDepth];
    uint32_t second_mask;     // bit d set: second operand of mix d
    uint32_t pad;
};

struct Material {
    uint32_t leaf_offset;
    uint32_t leaf_count;
    uint32_t param_offset;
    uint32_t emission_sampling;
    uint32_t may_emit;
    uint32_t program;  // generated surface program: one per distinct closure-leaf layout
    uint32_t pad[2];
};

// ---------------------------------------------------------------------------
// lights
// ---------------------------------------------------------------------------
struct Emitter {
    uint32_t instance;
    uint32_t prim;
    uint32_t material;
    uint32_t emission_sampling;
};

// Analytic lamp (LightGpu, src/luisa/path_tracer_types.h:142-168): point,
// spot or distant.  Axes are the normalised columns of the lamp transform,
// axis_scale their lengths.
struct Lamp {
    float position[3];
    uint32_t type;  // psx_light_type
    float axis_x[3];
    float power;
    float axis_y[3];
    float radius;
    float axis_z[3];
    float spot_angle;
    float axis_scale[3];
    float spot_smooth;
    float color[3];
    float angle;
    uint32_t flags;  // kLamp* bits
    uint32_t visibility;
    uint32_t max_bounces;
    float size_u;    // area lamps: edge lengths along axis_x and axis_y, scaled
    float size_v;
    float spread;    // area lamps: emission spread angle
    uint32_t pad0, pad1;
};
constexpr uint32_t kLampNormalize = 1u, kLampEllipse = 2u, kLampSphere = 4u, kLampUseMis = 8u, kLampForward = 32u,
                   kLampFullSpread = 64u, kLampCastShadow = 128u;
// lamp_bits: visibility mask in bits 0-7, forward-intersectable (not distant,
// non-zero size) in bit 8, effective MIS in bit 9
constexpr uint32_t kLampBitIntersectable = 256u, kLampBitMis = 512u;
constexpr uint32_t kLampPoint = 0u, kLampSpot = 1u, kLampArea = 2u, kLampDistant = 3u;

// Light tree records (LightTreeNodeGpu / LightTreeEmitterGpu,
// src/luisa/path_tracer_types.h:202-225): measure packed as bounds min +
// energy, bounds max + theta_o, cone axis + theta_e.
constexpr uint32_t kTreeHasBounds = 1u, kTreeHasOrientation = 2u, kTreeDistant = 4u;
constexpr uint32_t kTreeInner = 0u, kTreeLeaf = 1u, kTreeDistantNode = 2u;

// `trig` holds cos(theta_o), sin(theta_o), cos(theta_o + theta_e) and
// cos(theta_e - theta_o) of the measure, evaluated once on the host.
struct TreeNode {
    float4 bounds_min_energy;
    float4 bounds_max_theta_o;
    float4 cone_axis_theta_e;
    float4 trig;
    uint4 topology;  // parent, left, right, kind
    uint4 emitters;  // first, count, measure flags, 0
};

struct TreeEmitter {
    float4 bounds_min_energy;
    float4 bounds_max_theta_o;
    float4 cone_axis_theta_e;
    float4 trig;
    uint4 identity;  // emitter id, measure flags, 0, 0
};

// Node of the binary BVH over lamp bounding spheres that the forward-hit
// search walks: depth-first order, the left child follows its parent, and
// `miss` is the node after the subtree, so no stack is needed.  `leaf`:
// 0 for an inner node, else first | count << 24 into SceneView::lamp_order.
struct LampNode {
    float lo[3], hi[3];
    uint32_t miss, leaf;
};
constexpr uint32_t kLampBvhMinLamps = 9u;  // scenes with fewer lamps scan them in order

struct DistEntry {
    float cumulative;
    float selection_pdf;
    uint32_t kind;   // 0 triangle, 1 analytic, 2 environment, 3 sentinel
    uint32_t index;
};

// ---------------------------------------------------------------------------
// scene view passed to kernels (device pointers)
// ---------------------------------------------------------------------------
struct SceneView {
    const QNode4 *nodes;
    const LeafTri *tris;
    const Instance *instances;
    const Mesh *meshes;
    const float *positions;         // float3 array (object + applied world)
    const uint32_t *indices;        // uint3 array
    const float *normals;           // float3 array
    const uint8_t *smooth;
    const uint32_t *slots;          // mesh slot -> material
    const uint32_t *tri_slots;
    const uint32_t *overrides;
    const Material *materials;
    const Leaf *leaves;
    const float *params;
    const Emitter *emitters;
    const DistEntry *distribution;
    const float4 *material_shadow;    // per material: transparent extinction for shadow rays
    const float4 *material_emission;  // per material: constant emission
    const uint32_t *instance_opaque;  // per instance: 1 when every material it can use stops shadow rays
    const Lamp *lamps;
    const float4 *lamp_sphere;        // per lamp: centre, radius (forward-hit culling)
    const uint32_t *lamp_bits;
    const LampNode *lamp_nodes;       // forward-hit BVH over the lamps, or null
    const uint32_t *lamp_order;       // lamp indices of the BVH leaves
    uint32_t lamp_node_count;
    uint32_t lamp_count;
    float light_selection_pdf;        // flat-CDF probability of one lamp
    const TreeNode *tree_nodes;
    const TreeEmitter *tree_emitters;
    const uint2 *tree_mapping;        // emitter id -> (tree emitter, leaf node)
    const uint32_t *emitter_cprim;    // Cycles primitive of each emissive triangle
    uint32_t tree_node_count, tree_emitter_count, tree_root;
    uint32_t tlas_root;
    uint32_t instance_count;
    uint32_t material_count;
    uint32_t emitter_count;
    uint32_t distribution_count;
    float triangle_area_pdf;
    float background[3];
    uint32_t world_visibility;
    uint32_t has_tlas;
};

// ---------------------------------------------------------------------------
// per-render kernel parameters (RenderKernelParameters,
// src/luisa/path_tracer_types.h:409-448, restricted)
// ---------------------------------------------------------------------------
struct KernelParams {
    uint32_t window_x, window_y, window_width, window_height;
    uint32_t full_width, full_height, seed, sequence_size;
    uint32_t max_bounces, min_bounces, max_diffuse, max_glossy;
    uint32_t max_transmission, transparent_min, transparent_max, max_path_steps;
    uint32_t transparent_background, nee, mis_mode, use_tree;
    float clamp_direct, clamp_indirect, light_inv_rr_threshold, pass_alpha_threshold;
    float tan_h, tan_v, shift_x, shift_y;
    uint32_t projection;  // 0 perspective, 1 orthographic
    float ortho_span;     // vertical view-plane span of the orthographic camera
    float near_clip, far_clip, pad1, pad2;
    float camera[16];
};

// Light-pass sums per pixel, in the reference's LightPassBuffer order
// (src/luisa/path_tracer_internal.h:410-424).
constexpr uint32_t kLightPasses = 12u;
enum LightPassSlot : uint32_t {
    kLpDiffuseDirect, kLpDiffuseIndirect, kLpGlossyDirect, kLpGlossyIndirect, kLpTransmissionDirect,
    kLpTransmissionIndirect, kLpVolumeDirect, kLpVolumeIndirect, kLpEmission, kLpEnvironment, kLpGlossyColor,
    kLpTransmissionColor
};

struct FilmView {
    float *combined;   // float4 per pixel
    float *normal;     // float4 per pixel
    float *albedo;     // float4 per pixel
    uint32_t *sample_count;
    float *light;      // kLightPasses float4 per pixel, or null
};

struct Counters {
    unsigned long long rays;
    unsigned long long nodes;
    unsigned long long tris;
    unsigned long long pad;
};

// ---------------------------------------------------------------------------
// wavefront integrator: per-path state in HBM, structure of arrays indexed by
// path = sample_in_batch * padded_pixels + tile * 64 + lane (tiles of 64
// pixels, 8x8 unless the view's tile_shift says otherwise)
// ---------------------------------------------------------------------------
// trace_count and shade_count share one 8-byte word, so that a shade wave
// appends to both with a single atomic
struct alignas(16) WavefrontCounters {
    uint32_t trace_count;   // entries in this iteration's trace queue
    uint32_t shade_count;   // paths in this iteration's shade list
    uint32_t trace_fetch;   // work-fetch cursor of the persistent trace kernel
    uint32_t pad;
    uint32_t segment_fetch[8];  // per-segment cursors when the queue is split by XCD
};

// One request's share of a server batch: samples [first, first + count) of
// every pixel of its window, held at paths
// [path_base, path_base + count * padded) in window-tile order, and
// accumulated by lanes [pixel_base, pixel_base + padded).  `padded`,
// `path_base` and `pixel_base` are multiples of 64, so a wave never spans
// two items.
struct BatchItem {
    KernelParams k;
    FilmView film;
    const float *filter;
    const float4 *sobol;
    uint32_t first, count, padded, path_base;
    uint32_t pixel_base, pad0, pad1, pad2;
};

struct WavefrontView {
    float4 *ray_o;       // origin.xyz, tmin
    float4 *ray_d;       // direction.xyz, tmax
    uint4 *aux;          // source object, source primitive, step, bits
    float4 *throughput;  // xyz, previous BSDF pdf
    float4 *mis_normal;  // shading normal of the last non-transparent scatter
    float4 *radiance;    // xyz, transparency
    float4 *normal;      // primary normal AOV
    float4 *albedo;      // primary albedo AOV
    uint4 *state0;       // flag, visibility, bounce, diffuse bounce
    uint4 *state1;       // glossy, transmission, transparent bounce, rng offset
    uint2 *ident;        // sample index, pixel rng hash
    uint4 *hit;          // instance (kNone = miss), primitive, Cycles primitive, t bits
    float2 *hit_uv;
    float4 *shadow_o;    // origin.xyz, tmax
    float4 *shadow_d;    // direction.xyz, clamp depth bits
    uint4 *shadow_q;     // source object, source primitive, light object, light primitive
    float4 *shadow_l;    // light estimate xyz, transparent depth bits
    float4 *shadow_t;    // transmittance xyz
    uint32_t *queue[2];  // trace entries: path << 1 | is_shadow
    uint32_t *list[2];   // shade entries: path
    WavefrontCounters *counters;  // one per iteration
    float4 *lp;            // light passes: 10 float4 per path (LightSums), field-major, or null
    float4 *shadow_split;  // light passes: 2 float4 per path (shadow dw, gw), field-major, or null
    uint32_t capacity;            // paths
    uint32_t padded_pixels;       // 64 * tiles of the window
    // launch shape (any values in range give the same film):
    uint32_t tile_shift;    // a wave's 64 paths cover (1 << tile_shift) x (64 >> tile_shift) pixels
    uint32_t refill;        // idle lanes of a trace wave that trigger a refill, 1..64
    uint32_t trace_chunk;   // queue entries a trace wave reserves per atomic, >= 1
    uint32_t queue_chunk;   // next-queue slots a shade wave reserves per atomic, 128..kMaxQueueChunk
    uint32_t list_chunk;    // shade-list slots a shade wave reserves per atomic, 64..kMaxListChunk
    uint32_t shade_sort;    // 1: each shade block takes its 256 paths sorted by shade_class
    uint32_t segments;      // 1, or 8: the trace queue is split in 8 segments, one per XCD
    uint32_t pad1[3];
};

// Largest shade reservations; the queues' headroom is sized for them.
constexpr uint32_t kMaxQueueChunk = 256u, kMaxListChunk = 256u;

}  // namespace psx
