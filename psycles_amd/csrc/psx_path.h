// psx_path.h - device code of the MI355X render path shared by the
// megakernel and the wavefront integrator (gfx950).
//
// Camera rays, two-level BVH4 traversal with an LDS short stack and the exact
// Cycles closest-hit contract, closure-tree shading, emissive-triangle NEE
// with transparent shadows, Russian roulette and the path step.  Each stage
// cites the reference file whose behaviour it reproduces.
//
// Floating point: compiled with -ffp-contract=off and correctly rounded
// division/sqrt; fmaf appears exactly where the reference fuses.  Helper
// conventions (bit patterns fixed by this implementation):
//   dot(a, b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(v) = v * (1 / sqrt(dot(v, v)))
//   M * (v, w)   = ((c0*x + c1*y) + c2*z) + c3*w
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#endif

#include "psx_device.h"

#pragma clang fp contract(off)

namespace psx {
namespace {

// ===========================================================================
// vector algebra
// ===========================================================================
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 V(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ f3 neg(f3 a) { return V(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) {
    const float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
__device__ __forceinline__ f3 sel(f3 a, f3 b, bool c) { return c ? b : a; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ f3 vmax(f3 a, f3 b) {
    return V(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z));
}
__device__ __forceinline__ f3 vclamp(f3 a, float lo, float hi) {
    return V(clampf(a.x, lo, hi), clampf(a.y, lo, hi), clampf(a.z, lo, hi));
}
__device__ __forceinline__ bool any_nonzero(f3 a) { return a.x != 0.0f || a.y != 0.0f || a.z != 0.0f; }
__device__ __forceinline__ bool any_nan(f3 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z); }
__device__ __forceinline__ f3 load3(const float *p) { return V(p[0], p[1], p[2]); }

// safe_normalize (src/luisa/path_tracer_lighting.cpp:9-18)
__device__ __forceinline__ f3 safe_normalize(f3 value, f3 fallback) {
    const bool valid = dot(value, value) > 1.0e-20f;
    f3 s = sel(fallback, value, valid);
    const bool fallback_valid = dot(s, s) > 1.0e-20f;
    s = sel(V(0.0f, 0.0f, 1.0f), s, fallback_valid);
    return normalize(s);
}

// normalize_or (include/psycles/luisa/spherical_geometry.h:34-40)
__device__ __forceinline__ f3 normalize_or(f3 value, f3 fallback) {
    const float l2 = dot(value, value);
    return l2 > 1.0e-20f ? value / sqrtf(fmaxf(l2, 1.0e-20f)) : fallback;
}

// cycles_transform::point / direction (src/luisa/cycles_transform.cpp:6-27)
__device__ __forceinline__ f3 affine_point(const Affine &t, f3 v) {
    const float *e = t.m;
    return V(fmaf(v.x, e[0], fmaf(v.y, e[3], fmaf(v.z, e[6], e[9]))),
             fmaf(v.x, e[1], fmaf(v.y, e[4], fmaf(v.z, e[7], e[10]))),
             fmaf(v.x, e[2], fmaf(v.y, e[5], fmaf(v.z, e[8], e[11]))));
}
__device__ __forceinline__ f3 affine_dir(const Affine &t, f3 v) {
    const float *e = t.m;
    return V(fmaf(v.x, e[0], fmaf(v.y, e[3], fmaf(v.z, e[6], 0.0f))),
             fmaf(v.x, e[1], fmaf(v.y, e[4], fmaf(v.z, e[7], 0.0f))),
             fmaf(v.x, e[2], fmaf(v.y, e[5], fmaf(v.z, e[8], 0.0f))));
}
// Luisa float4x4 * float4 with w in {0, 1}
__device__ __forceinline__ f3 mat_mul(const Affine &t, f3 v, float w) {
    const float *e = t.m;
    return V(((e[0] * v.x + e[3] * v.y) + e[6] * v.z) + e[9] * w,
             ((e[1] * v.x + e[4] * v.y) + e[7] * v.z) + e[10] * w,
             ((e[2] * v.x + e[5] * v.y) + e[8] * v.z) + e[11] * w);
}
__device__ __forceinline__ f3 cam_point(const float *m, f3 v) {
    return V(fmaf(v.x, m[0], fmaf(v.y, m[4], fmaf(v.z, m[8], m[12]))),
             fmaf(v.x, m[1], fmaf(v.y, m[5], fmaf(v.z, m[9], m[13]))),
             fmaf(v.x, m[2], fmaf(v.y, m[6], fmaf(v.z, m[10], m[14]))));
}
__device__ __forceinline__ f3 cam_dir(const float *m, f3 v) {
    return V(fmaf(v.x, m[0], fmaf(v.y, m[4], fmaf(v.z, m[8], 0.0f))),
             fmaf(v.x, m[1], fmaf(v.y, m[5], fmaf(v.z, m[9], 0.0f))),
             fmaf(v.x, m[2], fmaf(v.y, m[6], fmaf(v.z, m[10], 0.0f))));
}

// ===========================================================================
// sampler (include/psycles/luisa/cycles_sampler.h:36-271)
// ===========================================================================
__device__ __forceinline__ uint32_t reverse_bits(uint32_t v) { return __brev(v); }

__device__ __forceinline__ uint32_t reversed_bit_owen(uint32_t v, uint32_t seed) {
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16u) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return v;
}

__device__ __forceinline__ uint32_t hash_wang_seeded(uint32_t v, uint32_t seed) {
    v = (v ^ 61u) ^ seed;
    v += v << 3u;
    v ^= v >> 4u;
    v *= 0x27d4eb2du;
    return v;
}

__device__ __forceinline__ uint32_t pattern_index(uint32_t dimension, uint32_t seed) {
    uint32_t v = dimension & 0xffu;
    v ^= seed;
    v *= 0xe170893du;
    v ^= seed >> 16u;
    v ^= (v & 0xffu) >> 4u;
    v ^= seed >> 8u;
    v *= 0x0929eb3fu;
    v ^= seed >> 23u;
    v ^= (v & 0xffu) >> 1u;
    v *= 1u | (seed >> 27u);
    v *= 0x6935fa69u;
    v ^= (v & 0xffu) >> 11u;
    v *= 0x74dcb303u;
    v ^= (v & 0xffu) >> 2u;
    v *= 0x9e501cc3u;
    v ^= (v & 0xffu) >> 2u;
    v *= 0xc860a3dfu;
    v &= 0xffu;
    v ^= v >> 5u;
    return v;
}

__device__ __forceinline__ uint32_t pixel_hash(uint32_t x, uint32_t y, uint32_t seed) {
    const uint32_t qx = 1103515245u * ((x >> 1u) ^ y);
    const uint32_t qy = 1103515245u * ((y >> 1u) ^ x);
    const uint32_t h = 1103515245u * (qx ^ (qy >> 3u));
    return h ^ seed;
}

__device__ __forceinline__ uint32_t shuffled_index(uint32_t sample, uint32_t dimension, uint32_t seed,
                                                   uint32_t size) {
    const uint32_t pattern = pattern_index(dimension, seed);
    const uint32_t mask = size - 1u;
    const uint32_t shuffled =
        reverse_bits(reversed_bit_owen(reverse_bits(sample), hash_wang_seeded(dimension, seed)));
    sample = (sample & ~mask) | (shuffled & mask);
    return ((pattern * size) + sample) % (size * 256u);
}

__device__ __forceinline__ float4 sample4(const float4 *table, uint32_t size, uint32_t sample,
                                          uint32_t hash, uint32_t dimension) {
    return table[shuffled_index(sample, dimension, hash, size)];
}

// ===========================================================================
// path state constants (include/psycles/luisa/cycles_path_state.h)
// ===========================================================================
constexpr uint32_t FLAG_REFLECT = 1u << 0, FLAG_SINGULAR = 1u << 1, FLAG_TRANSPARENT = 1u << 2,
                   FLAG_IMPORTANCE_BAKE = 1u << 3, FLAG_DIFFUSE_ANCESTOR = 1u << 4,
                   FLAG_MIS_HAD_TRANSMISSION = 1u << 6, FLAG_MIS_SKIP = 1u << 7,
                   FLAG_SINGLE_PASS_DONE = 1u << 8, FLAG_TRANSPARENT_BACKGROUND = 1u << 9,
                   FLAG_TERMINATE_ON_NEXT_SURFACE = 1u << 10,
                   FLAG_TERMINATE_AFTER_TRANSPARENT = 1u << 12, FLAG_SURFACE_PASS = 1u << 18,
                   FLAG_ANY_PASS = (1u << 18) | (1u << 19);
constexpr uint32_t CVIS_CAMERA = 1u, CVIS_TRANSMIT = 2u, CVIS_DIFFUSE = 4u, CVIS_GLOSSY = 8u,
                   CVIS_VOLUME_SCATTER = 16u;
constexpr uint32_t LABEL_TRANSMIT = 1u, LABEL_REFLECT = 2u, LABEL_DIFFUSE = 4u, LABEL_GLOSSY = 8u,
                   LABEL_SINGULAR = 16u, LABEL_TRANSPARENT = 32u, LABEL_TRANSMIT_TRANSPARENT = 128u,
                   LABEL_RAY_PORTAL = 512u;
constexpr uint32_t RT_BACKFACING = 1u, RT_BSDF = 8u, RT_BSDF_HAS_EVAL = 16u, RT_TRANSPARENT = 1024u,
                   RT_BSDF_HAS_TRANSMISSION = 2048u, RT_RAY_PORTAL = 4096u;
constexpr uint32_t EV_DIFFUSE = 1u, EV_GLOSSY = 2u, EV_SINGULAR = 4u, EV_REFLECTION = 8u,
                   EV_TRANSMISSION = 16u, EV_TRANSPARENT = 32u;
constexpr uint32_t VIS_CAMERA = 1u, VIS_DIFFUSE = 2u, VIS_GLOSSY = 4u, VIS_TRANSMISSION = 8u,
                   VIS_SHADOW = 16u, VIS_VOLUME_SCATTER = 32u;
constexpr uint32_t MESH_LIGHT_VISIBILITY = VIS_DIFFUSE | VIS_GLOSSY | VIS_TRANSMISSION | VIS_VOLUME_SCATTER;

struct PathState {
    uint32_t flag, visibility, bounce, diffuse_bounce, glossy_bounce, transmission_bounce,
        transparent_bounce, rng_offset;
};

// next_surface (cycles_path_state.h:205-362)
__device__ __forceinline__ void next_surface(PathState &s, uint32_t label, uint32_t runtime_flag,
                                             const KernelParams &k) {
    const bool transparent = (label & (LABEL_TRANSPARENT | LABEL_RAY_PORTAL)) != 0u;
    if (transparent) {
        s.transparent_bounce += 1u;
        s.rng_offset += 16u;
        s.flag |= FLAG_TRANSPARENT;
        if (s.transparent_bounce >= k.transparent_max) s.flag |= FLAG_TERMINATE_ON_NEXT_SURFACE;
        if ((runtime_flag & RT_RAY_PORTAL) != 0u) s.flag |= FLAG_MIS_SKIP;
        return;
    }
    s.bounce += 1u;
    s.rng_offset += 16u;
    if (s.bounce >= k.max_bounces) s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    s.visibility = 0u;
    s.flag &= ~(FLAG_REFLECT | FLAG_SINGULAR | FLAG_TRANSPARENT | FLAG_IMPORTANCE_BAKE | FLAG_MIS_SKIP |
                FLAG_MIS_HAD_TRANSMISSION);
    const bool reflection = (label & LABEL_REFLECT) != 0u;
    const bool diffuse = (label & LABEL_DIFFUSE) != 0u;
    const bool glossy = (label & LABEL_GLOSSY) != 0u;
    const bool transmit_transparent = (label & LABEL_TRANSMIT_TRANSPARENT) != 0u;
    if (reflection) s.flag |= FLAG_REFLECT;
    if (reflection || !transmit_transparent) s.flag &= ~FLAG_TRANSPARENT_BACKGROUND;
    const bool diffuse_reflection = reflection && diffuse;
    const bool glossy_reflection = reflection && !diffuse;
    const bool transmission = !reflection;
    if (diffuse_reflection) s.diffuse_bounce += 1u;
    if (glossy_reflection) s.glossy_bounce += 1u;
    if (transmission) s.transmission_bounce += 1u;
    if (diffuse_reflection && s.diffuse_bounce >= k.max_diffuse) s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    if (glossy_reflection && s.glossy_bounce >= k.max_glossy) s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    if (transmission && s.transmission_bounce >= k.max_transmission)
        s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    if (transmission) s.visibility |= CVIS_TRANSMIT;
    s.visibility |= diffuse ? CVIS_DIFFUSE : CVIS_GLOSSY;
    if (diffuse) s.flag |= FLAG_DIFFUSE_ANCESTOR;
    if (!diffuse && !glossy) s.flag |= FLAG_SINGULAR | FLAG_MIS_SKIP;
    if ((runtime_flag & RT_BSDF_HAS_TRANSMISSION) != 0u) s.flag |= FLAG_MIS_HAD_TRANSMISSION;
    if ((s.flag & FLAG_ANY_PASS) == 0u && (s.flag & FLAG_TRANSPARENT_BACKGROUND) == 0u)
        s.flag |= FLAG_SURFACE_PASS;
}

// contract_visibility (cycles_path_state.h:448-489)
__device__ __forceinline__ uint32_t contract_visibility(uint32_t v) {
    const uint32_t t = (v & CVIS_TRANSMIT) ? (v & ~(CVIS_DIFFUSE | CVIS_GLOSSY)) : v;
    uint32_t r = 0u;
    if (t & CVIS_CAMERA) r |= VIS_CAMERA;
    if (t & CVIS_TRANSMIT) r |= VIS_TRANSMISSION;
    if (t & CVIS_DIFFUSE) r |= VIS_DIFFUSE;
    if (t & CVIS_GLOSSY) r |= VIS_GLOSSY;
    if (t & CVIS_VOLUME_SCATTER) r |= VIS_VOLUME_SCATTER;
    return r;
}

__device__ __forceinline__ uint32_t label_from_events(uint32_t e) {
    uint32_t l = 0u;
    if (e & EV_TRANSMISSION) l |= LABEL_TRANSMIT;
    if (e & EV_REFLECTION) l |= LABEL_REFLECT;
    if (e & EV_DIFFUSE) l |= LABEL_DIFFUSE;
    if (e & EV_GLOSSY) l |= LABEL_GLOSSY;
    if (e & EV_SINGULAR) l |= LABEL_SINGULAR;
    if (e & EV_TRANSPARENT) l |= LABEL_TRANSPARENT;
    return l;
}

// ===========================================================================
// traversal: two-level BVH4 with an LDS short stack, exact Cycles closest-hit
// contract (src/luisa/path_kernel_scene_traversal.cpp:34-68, 212-493;
//  src/luisa/cycles_triangle_intersection_component.cpp:14-108)
// ===========================================================================
constexpr int kStackDepth = static_cast<int>(kTraversalStackDepth);
constexpr uint32_t kShortStack = 16u;  // LDS entries per lane; deeper entries spill to HBM

struct Ray {
    f3 o, d;
    float tmin, tmax;
};

struct Hit {
    uint32_t inst, prim, object, cprim;
    float t, u, v;
    bool valid;
};

// Per-lane traversal stack: the top kShortStack entries live in LDS laid out
// [entry][lane] (bank-conflict free), older entries move to a per-lane column
// of an HBM spill buffer laid out [entry][lane of the grid].  Scene
// compilation bounds the depth by kTraversalStackDepth.
struct Stack {
    uint32_t *lds;    // block base of the LDS ring
    uint32_t *spill;  // this lane's spill column
    uint32_t lanes;   // LDS row stride (lanes per block)
    uint32_t lane;    // lane within the block
    uint32_t stride;  // spill row stride (lanes of the grid)
    int sp, bottom;   // entries [bottom, sp) are in LDS, [0, bottom) spilled

    __device__ __forceinline__ void push(uint32_t e) {
        if (sp >= kStackDepth) return;
        if (sp - bottom == static_cast<int>(kShortStack)) {
            spill[static_cast<size_t>(bottom) * stride] = lds[(bottom & (kShortStack - 1u)) * lanes + lane];
            ++bottom;
        }
        lds[(sp & (kShortStack - 1u)) * lanes + lane] = e;
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        // the LDS slot is read unconditionally (the index is always in the
        // ring), so the common case stays an LDS access instead of one
        // generic load that selects between LDS and the spill column
        uint32_t e = lds[(sp & (kShortStack - 1u)) * lanes + lane];
        if (sp < bottom) {
            bottom = sp;
            e = spill[static_cast<size_t>(sp) * stride];
        }
        return e;
    }
};

__device__ __forceinline__ Stack make_stack(uint32_t *lds, uint32_t *spill, uint32_t lanes, uint32_t lane,
                                            uint32_t global_lane, uint32_t stride) {
    Stack st;
    st.lds = lds;
    st.spill = spill + global_lane;
    st.lanes = lanes;
    st.lane = lane;
    st.stride = stride;
    st.sp = 0;
    st.bottom = 0;
    return st;
}

__device__ __forceinline__ float clamp_bvh_dir(float v) {
    const float m = 8.271806e-25f;
    return fabsf(v) > m ? v : copysignf(m, v);
}

__device__ __forceinline__ f3 pluecker_cross(f3 a, f3 b) {
    return V(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float pluecker_dot(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// d is already clamped with clamp_bvh_dir
__device__ __forceinline__ bool pluecker(f3 o, f3 d, float tmin, float tmax, f3 p0, f3 p1, f3 p2,
                                         float &t, float &bu, float &bv) {
    const f3 v0 = p0 - o, v1 = p1 - o, v2 = p2 - o;
    const f3 e0 = v2 - v0, e1 = v0 - v1, e2 = v1 - v2;
    const float u = pluecker_dot(pluecker_cross(e0, v2 + v0), d);
    const float v = pluecker_dot(pluecker_cross(e1, v0 + v1), d);
    const float w = pluecker_dot(pluecker_cross(e2, v1 + v2), d);
    const float sum = u + v + w;
    const float tol = FLT_EPSILON * fabsf(sum);
    const float mn = fminf(u, fminf(v, w));
    const float mx = fmaxf(u, fmaxf(v, w));
    const bool inside = (mn >= -tol) || (mx <= tol);
    const f3 n2 = pluecker_cross(e1, e0) * 2.0f;
    const float den = pluecker_dot(n2, d);
    const float safe_den = den != 0.0f ? den : 1.0f;
    t = pluecker_dot(v0, n2) / safe_den;
    const bool in_range = (t >= tmin) && (t <= tmax);
    const float rs = fabsf(sum) >= 1.0e-18f ? 1.0f / sum : 0.0f;
    bu = fminf(u * rs, 1.0f);
    bv = fminf(v * rs, 1.0f);
    return inside && den != 0.0f && in_range;
}

struct TraceQuery {
    uint32_t visibility;
    uint32_t src_obj, src_prim, light_obj, light_prim;
};

__device__ __forceinline__ bool order_accepts(const Hit &best, float t, uint32_t object, uint32_t prim) {
    if (!best.valid) return true;
    if (t < best.t) return true;
    if (t == best.t) return object > best.object || (object == best.object && prim > best.cprim);
    return false;
}

// compare-exchange of a sorting network on (distance, entry) pairs
__device__ __forceinline__ void sort_pair(float &ka, uint32_t &ea, float &kb, uint32_t &eb) {
    const bool swap = kb < ka;
    const float k0 = swap ? kb : ka, k1 = swap ? ka : kb;
    const uint32_t e0 = swap ? eb : ea, e1 = swap ? ea : eb;
    ka = k0;
    kb = k1;
    ea = e0;
    eb = e1;
}

// Material of a triangle of an instance: per-triangle slot, instance
// override, then the mesh's slot table (clamped to its last slot).
__device__ __forceinline__ uint32_t material_of(const SceneView &s, const Instance &inst, const Mesh &mesh,
                                                uint32_t prim) {
    const uint32_t slot = mesh.tri_slot_offset != kNone ? s.tri_slots[mesh.tri_slot_offset + prim] : 0u;
    if (slot < inst.override_count) return s.overrides[inst.override_offset + slot];
    if (mesh.slot_count == 0u) return kNone;
    return s.slots[mesh.slot_offset + (slot < mesh.slot_count ? slot : mesh.slot_count - 1u)];
}

__device__ __forceinline__ uint32_t hit_material(const SceneView &s, uint32_t inst_index, uint32_t prim) {
    const Instance &inst = s.instances[inst_index];
    const Mesh &mesh = s.meshes[inst.mesh];
    return material_of(s, inst, mesh, prim);
}

// Closest hit along [tmin, tmax] under the Cycles order: smallest t, ties to
// the greatest (object, primitive).  Box culling is conservative (boxes are
// padded at build time and the far distance is widened), so traversal order
// never changes the result, only its cost.
//
// The traversal is a resumable state machine (begin, then step until it
// reports completion) so that persistent kernels can refill lanes whose ray
// has finished while the rest of the wave keeps traversing.
struct Traversal {
    Ray ray;        // world ray of the query
    TraceQuery q;
    f3 o, d, inv, oi;  // ray in the current intersection space; oi = o * inv
    float lower;
    Hit best;
    uint32_t current;
    uint32_t inst_index, inst_prim_offset;
    bool in_blas;
    // Any-hit traversal (traversal_step<COUNT, true>, for shadow queries)
    // ends at the first accepted candidate on an instance whose every
    // material is opaque to shadow rays (`stop` while inside such an
    // instance; `opaque` once ended): the transparent shadow loop's
    // transmittance is then exactly zero whatever else lies on the ray, so
    // the early end changes no result bit.
    bool stop, opaque;
};

__device__ __forceinline__ f3 clamped_dir(f3 d) {
    return V(clamp_bvh_dir(d.x), clamp_bvh_dir(d.y), clamp_bvh_dir(d.z));
}

__device__ __forceinline__ void traversal_world_space(Traversal &t) {
    t.in_blas = false;
    t.o = t.ray.o;
    t.d = clamped_dir(t.ray.d);
    t.inv = V(1.0f / t.d.x, 1.0f / t.d.y, 1.0f / t.d.z);
    t.oi = t.o * t.inv;  // slab distances are fma(bound, inv, -o * inv)
}

__device__ __forceinline__ void traversal_begin(const SceneView &s, Traversal &t, Stack &st) {
    t.best.valid = false;
    t.best.t = t.ray.tmax;
    t.best.inst = t.best.prim = t.best.object = t.best.cprim = 0xffffffffu;
    t.best.u = t.best.v = 0.0f;
    t.lower = t.ray.tmin > 0.0f ? t.ray.tmin * 0.99999f : t.ray.tmin - 1.0e-30f;
    t.inst_index = t.inst_prim_offset = 0u;
    t.stop = t.opaque = false;
    traversal_world_space(t);
    st.sp = 0;
    st.bottom = 0;
    t.current = s.has_tlas ? (s.tlas_root | kKindTlas) : kEmptySlot;
}

// One traversal iteration: an instance entry, a leaf or an inner node.
// Returns false once the query is complete (t.best holds the result).
template <bool COUNT, bool ANYHIT = false>
__device__ __forceinline__ bool traversal_step(const SceneView &s, Traversal &t, Stack &st, uint32_t &node_count,
                                               uint32_t &tri_count) {
    if (t.current == kEmptySlot) {
        if (st.sp == 0) return false;
        t.current = st.pop();
    }
    const uint32_t entry = t.current;
    t.current = kEmptySlot;
    const uint32_t kind = entry & kKindMask;
    const uint32_t index = entry & ~kKindMask;
    // back at the top level: world ray again
    if (t.in_blas && (kind == kKindTlas || kind == kKindInstance)) traversal_world_space(t);
    if (kind == kKindInstance) {
        // enter an instance: object-space ray with the Cycles FMA transform
        const Instance &inst = s.instances[index];
        if ((inst.visibility & t.q.visibility) == 0u) return true;
        t.inst_index = index;
        t.inst_prim_offset = inst.prim_offset;
        t.in_blas = true;
        if (ANYHIT) t.stop = s.instance_opaque[index] != 0u;
        if (!inst.applied) {
            t.o = affine_point(inst.w2o, t.ray.o);
            t.d = clamped_dir(affine_dir(inst.w2o, t.ray.d));
            t.inv = V(1.0f / t.d.x, 1.0f / t.d.y, 1.0f / t.d.z);
            t.oi = t.o * t.inv;
        }
        t.current = inst.blas_root | kKindBlas;
        return true;
    }
    // One 64-byte fetch serves both kinds of step: a node record, or a leaf
    // triangle (48 bytes, the next 16 read and ignored; the triangle array
    // carries one record of padding).  Lanes of a wave that sit on nodes and
    // on triangles then wait on one memory round trip per step, not two.
    const bool is_leaf = kind == kKindLeaf;
    const uint4 *rec = is_leaf ? reinterpret_cast<const uint4 *>(s.tris + (index >> 3))
                               : reinterpret_cast<const uint4 *>(s.nodes + index);
    const uint4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    if (is_leaf) {
        // one triangle per step: a wave whose lanes sit on leaves of
        // different sizes then runs one intersection test per step instead
        // of the longest leaf's loop; the rest of the leaf is the next step
        const uint32_t first = index >> 3;
        const uint32_t n = (index & 7u) + 1u;
        if (n > 1u) t.current = (((first + 1u) << 3) | (n - 2u)) | kKindLeaf;
        {
            if (COUNT) tri_count++;
            float th, bu, bv;
            const bool ok = pluecker(t.o, t.d, t.ray.tmin, t.ray.tmax,
                                     V(__uint_as_float(r0.x), __uint_as_float(r0.y), __uint_as_float(r0.z)),
                                     V(__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z)),
                                     V(__uint_as_float(r2.x), __uint_as_float(r2.y), __uint_as_float(r2.z)), th, bu, bv);
            if (!ok) return true;
            const uint32_t prim = r0.w;
            const uint32_t cprim = t.inst_prim_offset + prim;
            const bool excluded = (t.inst_index == t.q.src_obj && cprim == t.q.src_prim) ||
                                  (t.inst_index == t.q.light_obj && cprim == t.q.light_prim);
            if (excluded) return true;
            if (ANYHIT && t.stop) {
                // any opaque candidate in [tmin, tmax] zeroes the shadow
                t.opaque = true;
                t.best.valid = true;
                t.best.t = th;
                t.best.u = bu;
                t.best.v = bv;
                t.best.inst = t.best.object = t.inst_index;
                t.best.prim = prim;
                t.best.cprim = cprim;
                t.current = kEmptySlot;
                st.sp = 0;
                st.bottom = 0;
                return true;
            }
            if (!order_accepts(t.best, th, t.inst_index, cprim)) return true;
            t.best.valid = true;
            t.best.t = th;
            t.best.u = bu;
            t.best.v = bv;
            t.best.inst = t.inst_index;
            t.best.prim = prim;
            t.best.object = t.inst_index;
            t.best.cprim = cprim;
        }
        return true;
    }
    // inner node: test the four child boxes, continue with the nearest hit
    // child and push the others far-to-near
    const uint4 head = r0, qxy = r1, qz = r2, refs = r3;
    if (COUNT) node_count++;
    const float upper = (t.best.valid ? t.best.t : t.ray.tmax) * 1.00001f + 1.0e-30f;
    // slab distance of origin + q * 2^e is fma(q, 2^e * inv, origin * inv - o * inv)
    auto pow2 = [](uint32_t bits, int shift) {
        const int e = static_cast<int>(static_cast<int8_t>((bits >> shift) & 0xffu));
        return __uint_as_float(static_cast<uint32_t>(e + 127) << 23);
    };
    const float sx = pow2(head.w, 0) * t.inv.x, sy = pow2(head.w, 8) * t.inv.y, sz = pow2(head.w, 16) * t.inv.z;
    const float bx = fmaf(__uint_as_float(head.x), t.inv.x, -t.oi.x);
    const float by = fmaf(__uint_as_float(head.y), t.inv.y, -t.oi.y);
    const float bz = fmaf(__uint_as_float(head.z), t.inv.z, -t.oi.z);
    const uint32_t ref_arr[4] = {refs.x, refs.y, refs.z, refs.w};
    // Per axis, the plane nearer along the ray: the scale has the sign of
    // the inverse direction, lo <= hi, and fma rounds monotonically, so
    // min(t(lo), t(hi)) is exactly t(lo) for a positive scale and t(hi) for a
    // negative one.  Swapping the byte words once per node replaces a min and
    // a max per axis and child.
    const uint32_t nx = sx < 0.0f ? qxy.y : qxy.x, fx = sx < 0.0f ? qxy.x : qxy.y;
    const uint32_t ny = sy < 0.0f ? qxy.w : qxy.z, fy = sy < 0.0f ? qxy.z : qxy.w;
    const uint32_t nz = sz < 0.0f ? qz.y : qz.x, fz = sz < 0.0f ? qz.x : qz.y;
    float key[4];
    uint32_t ent[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        // conservative: boxes are padded by 1e-5 of their coordinate
        // magnitude before quantisation, which rounds outward, and `far` is
        // widened by 1e-5, all far above the rounding of these slab terms
        const int sh = 8 * i;
        const float t0x = fmaf(static_cast<float>((nx >> sh) & 0xffu), sx, bx);
        const float t1x = fmaf(static_cast<float>((fx >> sh) & 0xffu), sx, bx);
        const float t0y = fmaf(static_cast<float>((ny >> sh) & 0xffu), sy, by);
        const float t1y = fmaf(static_cast<float>((fy >> sh) & 0xffu), sy, by);
        const float t0z = fmaf(static_cast<float>((nz >> sh) & 0xffu), sz, bz);
        const float t1z = fmaf(static_cast<float>((fz >> sh) & 0xffu), sz, bz);
        const float near = fmaxf(fmaxf(t0x, t0y), t0z);
        const float far = fminf(fminf(t1x, t1y), t1z);
        const float far_safe = far * 1.00001f + 1.0e-30f;
        const uint32_t ref = ref_arr[i];
        const bool hit = ref != kEmptySlot && near <= far_safe && far_safe >= t.lower && near <= upper;
        key[i] = hit ? near : INFINITY;
        ent[i] = ref;
    }
    sort_pair(key[0], ent[0], key[1], ent[1]);
    sort_pair(key[2], ent[2], key[3], ent[3]);
    sort_pair(key[0], ent[0], key[2], ent[2]);
    sort_pair(key[1], ent[1], key[3], ent[3]);
    sort_pair(key[1], ent[1], key[2], ent[2]);
    if (key[3] != INFINITY) st.push(ent[3]);
    if (key[2] != INFINITY) st.push(ent[2]);
    if (key[1] != INFINITY) st.push(ent[1]);
    if (key[0] != INFINITY) t.current = ent[0];
    return true;
}

template <bool COUNT, bool ANYHIT = false>
__device__ Hit closest_hit(const SceneView &s, const Ray &ray, const TraceQuery &q, Stack &st,
                           uint32_t &node_count, uint32_t &tri_count, bool *opaque = nullptr) {
    Traversal t;
    t.ray = ray;
    t.q = q;
    traversal_begin(s, t, st);
    while (traversal_step<COUNT, ANYHIT>(s, t, st, node_count, tri_count)) {
    }
    if (opaque) *opaque = t.opaque;
    return t.best;
}

// ===========================================================================
// surface reconstruction (src/luisa/cycles_triangle_surface_component.cpp,
// src/luisa/path_kernel_surface_primitive.cpp:139-460)
// ===========================================================================
struct Surface {
    uint32_t inst, object, cprim, material, emission_sampling;
    bool applied, back_facing;
    f3 p0, p1, p2;      // intersection-space positions
    f3 wp0, wp1, wp2;   // world positions
    f3 position, object_hit_position, ng, ns, incoming;
};

__device__ __forceinline__ f3 interp3(float u, float v, f3 a, f3 b, f3 c) {
    return (a * (1.0f - u - v) + b * u) + c * v;
}

__device__ void resolve_surface(const SceneView &s, const Hit &h, const Ray &ray, Surface &sf) {
    const Instance &inst = s.instances[h.inst];
    const Mesh &mesh = s.meshes[inst.mesh];
    sf.inst = h.inst;
    sf.object = h.object;
    sf.cprim = h.cprim;
    sf.material = material_of(s, inst, mesh, h.prim);
    sf.applied = inst.applied != 0u;
    const uint32_t *tri = s.indices + 3u * (mesh.index_offset + h.prim);
    const uint32_t i0 = tri[0], i1 = tri[1], i2 = tri[2];
    const f3 op0 = load3(s.positions + 3u * (mesh.vertex_offset + i0));
    const f3 op1 = load3(s.positions + 3u * (mesh.vertex_offset + i1));
    const f3 op2 = load3(s.positions + 3u * (mesh.vertex_offset + i2));
    f3 p0 = op0, p1 = op1, p2 = op2;
    if (sf.applied) {
        p0 = load3(s.positions + 3u * (inst.ipos_offset + i0));
        p1 = load3(s.positions + 3u * (inst.ipos_offset + i1));
        p2 = load3(s.positions + 3u * (inst.ipos_offset + i2));
    }
    sf.p0 = p0;
    sf.p1 = p1;
    sf.p2 = p2;
    sf.wp0 = sf.applied ? p0 : affine_point(inst.o2w, op0);
    sf.wp1 = sf.applied ? p1 : affine_point(inst.o2w, op1);
    sf.wp2 = sf.applied ? p2 : affine_point(inst.o2w, op2);
    const f3 ogn = safe_normalize(cross(op1 - op0, op2 - op0), V(0.0f, 0.0f, 1.0f));
    const f3 tgn = safe_normalize(affine_dir(inst.n2w, ogn), neg(ray.d));
    f3 gn = tgn;
    if (sf.applied) {
        f3 sgn = safe_normalize(cross(p1 - p0, p2 - p0), tgn);
        if (dot(sgn, tgn) < 0.0f) sgn = neg(sgn);
        gn = sgn;
    }
    const bool smooth = mesh.smooth_offset != kNone && mesh.normal_offset != kNone &&
                        s.smooth[mesh.smooth_offset + h.prim] != 0u;
    f3 ns;
    if (smooth) {
        const f3 n0 = load3(s.normals + 3u * (mesh.normal_offset + i0));
        const f3 n1 = load3(s.normals + 3u * (mesh.normal_offset + i1));
        const f3 n2 = load3(s.normals + 3u * (mesh.normal_offset + i2));
        if (sf.applied) {
            const f3 w0 = safe_normalize(affine_dir(inst.n2w, n0), gn);
            const f3 w1 = safe_normalize(affine_dir(inst.n2w, n1), gn);
            const f3 w2 = safe_normalize(affine_dir(inst.n2w, n2), gn);
            ns = safe_normalize(interp3(h.u, h.v, w0, w1, w2), gn);
        } else {
            ns = safe_normalize(affine_dir(inst.n2w, interp3(h.u, h.v, n0, n1, n2)), gn);
        }
    } else {
        ns = sf.applied ? safe_normalize(gn, gn) : safe_normalize(affine_dir(inst.n2w, ogn), gn);
    }
    const bool back_facing = dot(gn, neg(ray.d)) < 0.0f;
    if (back_facing) {
        gn = neg(gn);
        ns = neg(ns);
    }
    if (dot(ns, gn) < 0.0f) ns = neg(ns);
    sf.back_facing = back_facing;
    sf.ng = gn;
    sf.ns = ns;
    if (sf.applied) {
        sf.position = (p0 + (p1 - p0) * h.u) + (p2 - p0) * h.v;
        sf.object_hit_position = sf.position;
    } else {
        const f3 object_position = (op0 + (op1 - op0) * h.u) + (op2 - op0) * h.v;
        sf.position = affine_point(inst.o2w, object_position);
        sf.object_hit_position = mat_mul(inst.w2o, sf.position, 1.0f);
    }
    sf.incoming = neg(ray.d);
    sf.emission_sampling = 0u;
    if (sf.material != kNone) {
        const Material &m = s.materials[sf.material];
        if (m.may_emit && (inst.visibility & MESH_LIGHT_VISIBILITY) != 0u) sf.emission_sampling = m.emission_sampling;
    }
}

// offset_ray_origin: Waechter & Binder, Ray Tracing Gems (2019), chapter 6.
__device__ __forceinline__ float offset_component(float p, float n) {
    const int32_t of_i = (int32_t)(256.0f * n);
    const float p_i = __int_as_float(__float_as_int(p) + (p < 0.0f ? -of_i : of_i));
    return fabsf(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : p_i;
}

// include/psycles/luisa/surface_ray.h:50-106
// source_triangle_covers_ray_origin: the three Pluecker edge functions share
// a strict sign, beyond a scaled epsilon, exactly when the ray from the
// origin crosses the source triangle's interior
__device__ __forceinline__ bool source_covers_origin(f3 oo, f3 od, f3 p0, f3 p1, f3 p2) {
    const f3 v0 = p0 - oo, v1 = p1 - oo, v2 = p2 - oo;
    const f3 e0 = v2 - v0, e1 = v0 - v1, e2 = v1 - v2;
    const float u = dot(cross(v2 + v0, e0), od);
    const float v = dot(cross(v0 + v1, e1), od);
    const float w = dot(cross(v1 + v2, e2), od);
    const float eps = FLT_EPSILON * fabsf(u + v + w);
    const float mn = fminf(u, fminf(v, w));
    const float mx = fmaxf(u, fmaxf(v, w));
    return (mn >= eps) || (mx <= -eps);
}

__device__ __forceinline__ f3 origin_with_self_exclusion(f3 wo, f3 wn, f3 oo, f3 od, f3 p0, f3 p1, f3 p2) {
    if (source_covers_origin(oo, od, p0, p1, p2)) return wo;
    return V(offset_component(wo.x, wn.x), offset_component(wo.y, wn.y), offset_component(wo.z, wn.z));
}

__device__ __forceinline__ f3 surface_ray_origin(const SceneView &s, const Surface &sf, f3 dir) {
    const Instance &inst = s.instances[sf.inst];
    const f3 co = sf.applied ? sf.position : sf.object_hit_position;
    const f3 cd = sf.applied ? dir : affine_dir(inst.w2o, dir);
    return origin_with_self_exclusion(sf.position, sf.ng, co, cd, sf.p0, sf.p1, sf.p2);
}

// Cycles' parabolic smooth-surface height (include/psycles/luisa/surface_ray.h:
// 108-181) in intersection space; the interpolated normal goes to world space
// with the instance's FMA affine unless the transform is applied.
__device__ f3 smooth_surface_offset(const Affine &o2w, bool applied, float bu, float bv, f3 p0, f3 p1, f3 p2,
                                    f3 n0, f3 n1, f3 n2, f3 world_ng) {
    const float u = 1.0f - bu - bv;
    const float v = bu;
    const float w = bv;
    const f3 object_position = (p0 * u + p1 * v) + p2 * w;
    const f3 object_normal = (n0 * u + n1 * v) + n2 * w;
    const f3 world_normal = applied ? object_normal : affine_dir(o2w, object_normal);
    const float a = dot(n2 - n0, p0 - p2);
    const float b = dot(n2 - n1, p1 - p2);
    const float c = dot(n1 - n0, p1 - p0);
    const float parabolic = ((a * u) * (u - 1.0f) + ((a + b) + c) * u * v) + (b * v) * (v - 1.0f);
    float f0 = fmaxf(fmaxf(dot(p1 - p0, n0), dot(p2 - p0, n0)), 0.0f);
    float f1 = fmaxf(fmaxf(dot(p0 - p1, n1), dot(p2 - p1, n1)), 0.0f);
    float f2 = fmaxf(fmaxf(dot(p0 - p2, n2), dot(p1 - p2, n2)), 0.0f);
    f0 = fmaxf(dot(p0 - object_position, n0) + f0, 0.0f);
    f1 = fmaxf(dot(p1 - object_position, n1) + f1, 0.0f);
    f2 = fmaxf(dot(p2 - object_position, n2) + f2, 0.0f);
    const float front = fmaxf(fminf(f0, fminf(f1, f2)), parabolic * 0.5f);
    float b0 = fmaxf(fmaxf(dot(p0 - p1, n0), dot(p0 - p2, n0)), 0.0f);
    float b1 = fmaxf(fmaxf(dot(p1 - p0, n1), dot(p1 - p2, n1)), 0.0f);
    float b2 = fmaxf(fmaxf(dot(p2 - p0, n2), dot(p2 - p1, n2)), 0.0f);
    b0 = fmaxf(dot(object_position - p0, n0) + b0, 0.0f);
    b1 = fmaxf(dot(object_position - p1, n1) + b1, 0.0f);
    b2 = fmaxf(dot(object_position - p2, n2) + b2, 0.0f);
    const float back = fminf(-fminf(b0, fminf(b1, b2)), parabolic * 0.5f);
    const float height = dot(world_normal, world_ng) > 0.0f ? front : back;
    return world_normal * height;
}

struct ShadowOrigin {
    f3 position;
    bool skip_self;  // the source primitive stays excluded from the shadow query
};

// shadow_terminator_origin (include/psycles/luisa/surface_ray.h:183-243)
__device__ ShadowOrigin shadow_terminator_origin(f3 position, f3 ns, f3 ng, f3 light_dir, float geometry_offset,
                                                 bool smooth, const Affine &o2w, bool applied, float bu, float bv,
                                                 f3 p0, f3 p1, f3 p2, f3 n0, f3 n1, f3 n2) {
    float normal_light = dot(ns, light_dir);
    const bool transmit = normal_light < 0.0f;
    normal_light = fabsf(normal_light);
    const f3 offset_normal = transmit ? neg(ng) : ng;
    const float geometric_light = dot(offset_normal, light_dir);
    const float safe_cutoff = fmaxf(geometry_offset, 1.0e-20f);
    const float near_terminator = clampf(2.0f - (geometric_light + normal_light) / safe_cutoff, 0.0f, 1.0f);
    const float regular = clampf(1.0f - geometric_light / safe_cutoff, 0.0f, 1.0f);
    const float amount = normal_light < geometry_offset ? near_terminator : regular;
    ShadowOrigin r{position, true};
    if (smooth && geometry_offset > 0.0f && amount > 0.0f) {
        const f3 offset = smooth_surface_offset(o2w, applied, bu, bv, p0, p1, p2, n0, n1, n2, offset_normal);
        r.position = position + offset * amount;
        r.skip_self = geometric_light > 0.0f;
    }
    return r;
}

// The lift, then the source-triangle certificate with the robust offset
// fallback while the source stays excluded (make_shadow_origin,
// src/luisa/path_kernel_surface_geometry.cpp:168-182; surface_shadow_origin,
// include/psycles/luisa/surface_ray.h:252-309).  Vertex normals are reloaded
// only for smooth triangles of instances with a terminator offset.
__device__ ShadowOrigin shadow_origin_core(f3 position, f3 ns, f3 ng, f3 dir, float geometry_offset, bool smooth,
                                           const Affine &o2w, const Affine &w2o, bool applied, float bu, float bv,
                                           f3 p0, f3 p1, f3 p2, f3 n0, f3 n1, f3 n2) {
    ShadowOrigin o = shadow_terminator_origin(position, ns, ng, dir, geometry_offset, smooth, o2w, applied, bu, bv,
                                              p0, p1, p2, n0, n1, n2);
    if (o.skip_self) {
        const f3 oo = applied ? o.position : affine_point(w2o, o.position);
        const f3 od = applied ? dir : affine_dir(w2o, dir);
        o.position = origin_with_self_exclusion(o.position, ng, oo, od, p0, p1, p2);
    }
    return o;
}

__device__ ShadowOrigin surface_shadow_origin(const SceneView &s, const Surface &sf, const Hit &h, f3 dir) {
    const Instance &inst = s.instances[sf.inst];
    const float offset = inst.terminator_offset;
    bool lift = false;
    f3 n0 = V(0.0f, 0.0f, 0.0f), n1 = n0, n2 = n0;
    if (offset > 0.0f) {
        const Mesh &mesh = s.meshes[inst.mesh];
        lift = mesh.smooth_offset != kNone && mesh.normal_offset != kNone && s.smooth[mesh.smooth_offset + h.prim] != 0u;
        if (lift) {
            const uint32_t *tri = s.indices + 3u * (mesh.index_offset + h.prim);
            n0 = load3(s.normals + 3u * (mesh.normal_offset + tri[0]));
            n1 = load3(s.normals + 3u * (mesh.normal_offset + tri[1]));
            n2 = load3(s.normals + 3u * (mesh.normal_offset + tri[2]));
            if (sf.applied) {
                // world vertex normals fall back to the unflipped geometric normal
                const f3 gn = sf.back_facing ? neg(sf.ng) : sf.ng;
                n0 = safe_normalize(affine_dir(inst.n2w, n0), gn);
                n1 = safe_normalize(affine_dir(inst.n2w, n1), gn);
                n2 = safe_normalize(affine_dir(inst.n2w, n2), gn);
            }
        }
    }
    return shadow_origin_core(sf.position, sf.ns, sf.ng, dir, offset, lift, inst.o2w, inst.w2o, sf.applied, h.u, h.v,
                              sf.p0, sf.p1, sf.p2, n0, n1, n2);
}

__device__ __forceinline__ float next_float_up(float t) {
    return t == 0.0f ? FLT_MIN : __uint_as_float(__float_as_uint(t) + 1u);
}
// ===========================================================================
// closures (src/luisa/graph_surface_closures.cpp:6-353,
// src/luisa/graph_surface_physical_closures.cpp:200-470,
// src/luisa/surface_closure_evaluation.cpp:67-464,
// src/luisa/surface_closure_sampling_core.cpp:96-676)
// ===========================================================================
constexpr uint32_t kMaxClosures = 4;
constexpr uint32_t KIND_DIFFUSE = 1u, KIND_TRANSPARENT = 2u;
constexpr float INV_PI = 0.31830988618379067154f;
constexpr float PI_F = 3.14159265358979323846f;

struct ClosureSet {
    uint32_t count;
    uint32_t kind[kMaxClosures];
    f3 weight[kMaxClosures];
    float sw[kMaxClosures];
    float roughness[kMaxClosures];
    f3 color[kMaxClosures];
    f3 normal;          // safe_normalize(Ns, Ns): every closure uses the shading normal
    f3 emission;
};

__device__ __forceinline__ float avg3(f3 v) { return (v.x + v.y + v.z) / 3.0f; }

#ifdef PSX_GENERATED_SURFACE_PROGRAMS
// Defined by the generated source compiled at run time: the closure set of
// surface program `program` over the parameter block `p`, statement for
// statement what the interpreter below computes for that leaf layout.
__device__ void generated_closures(uint32_t program, const float *p, ClosureSet &cs);
#endif

__device__ void material_closures(const SceneView &s, uint32_t material, f3 ns, ClosureSet &cs) {
    cs.count = 0u;
    cs.emission = V(0.0f, 0.0f, 0.0f);
    cs.normal = safe_normalize(ns, ns);
    if (material == kNone) return;
    const Material &m = s.materials[material];
    const float *p = s.params + m.param_offset;
#ifdef PSX_GENERATED_SURFACE_PROGRAMS
    generated_closures(m.program, p, cs);
    return;
#endif
    // the interpreter: one pass over the material's flattened closure leaves
    for (uint32_t l = 0; l < m.leaf_count; ++l) {
        const Leaf &leaf = s.leaves[m.leaf_offset + l];
        float w = 1.0f;
        for (uint32_t d = 0; d < leaf.depth; ++d) {
            const float f = clampf(p[leaf.factor[d]], 0.0f, 1.0f);
            w = w * (((leaf.second_mask >> d) & 1u) ? f : (1.0f - f));
        }
        const f3 color = V(p[leaf.color], p[leaf.color + 1u], p[leaf.color + 2u]);
        if (leaf.op == kLeafEmission) {
            cs.emission = cs.emission + (color * p[leaf.aux]) * w;
            continue;
        }
        // diffuse or transparent closure candidate
        const bool diffuse = leaf.op == kLeafDiffuse;
        if (!diffuse && leaf.op != kLeafTransparent) continue;
        f3 wt;
        float sw;
        if (diffuse) {
            wt = vmax(color * w, V(0.0f, 0.0f, 0.0f));
            if (!(avg3(wt) >= 1.0e-5f)) wt = V(0.0f, 0.0f, 0.0f);
            sw = fabsf(avg3(vmax(wt, V(0.0f, 0.0f, 0.0f))));
            wt = vmax(wt, V(0.0f, 0.0f, 0.0f));
        } else {
            wt = color * w;
            sw = fabsf(avg3(wt));
        }
        if (!(sw >= 1.0e-5f) || cs.count >= kMaxClosures) continue;
        const uint32_t kind = diffuse ? KIND_DIFFUSE : KIND_TRANSPARENT;
        const float roughness = diffuse ? p[leaf.aux] : 0.0f;
        const f3 ccolor = diffuse ? color : V(1.0f, 1.0f, 1.0f);
        // slot writes through constant indices keep the set in registers
#pragma unroll
        for (uint32_t c = 0; c < kMaxClosures; ++c) {
            if (c == cs.count) {
                cs.kind[c] = kind;
                cs.weight[c] = wt;
                cs.sw[c] = sw;
                cs.roughness[c] = roughness;
                cs.color[c] = ccolor;
            }
        }
        cs.count += 1u;
    }
}

__device__ __forceinline__ f3 material_constant_emission(const SceneView &s, uint32_t material) {
    f3 e = V(0.0f, 0.0f, 0.0f);
    const Material &m = s.materials[material];
    const float *p = s.params + m.param_offset;
    for (uint32_t l = 0; l < m.leaf_count; ++l) {
        const Leaf &leaf = s.leaves[m.leaf_offset + l];
        if (leaf.op != kLeafEmission) continue;
        float w = 1.0f;
        for (uint32_t d = 0; d < leaf.depth; ++d) {
            const float f = clampf(p[leaf.factor[d]], 0.0f, 1.0f);
            w = w * (((leaf.second_mask >> d) & 1u) ? f : (1.0f - f));
        }
        e = e + (V(p[leaf.color], p[leaf.color + 1u], p[leaf.color + 2u]) * p[leaf.aux]) * w;
    }
    return e;
}

__device__ __forceinline__ float oren_nayar_g(float cosine) {
    const float c = clampf(cosine, 0.0f, 1.0f);
    const float sine = sqrtf(fmaxf(1.0f - c * c, 0.0f));
    const float theta = acosf(c);
    const float safe_c = fmaxf(c, 1.0e-6f);
    const float regular =
        sine * (theta - 2.0f / 3.0f - sine * c) + 2.0f / 3.0f * (sine / safe_c) * (1.0f - sine * sine * sine);
    const float taylor = (PI_F * 0.5f - 2.0f / 3.0f) - c;
    return c < 1.0e-6f ? taylor : regular;
}

// diffuse_intensity (src/luisa/graph_surface_scattering.cpp:475-506)
__device__ __forceinline__ f3 diffuse_intensity(f3 n, float roughness, f3 color, f3 in, f3 out) {
    const float nl = fmaxf(dot(n, out), 0.0f);
    const float sigma = clampf(roughness, 0.0f, 1.0f);
    if (sigma < 1.0e-5f) {
        const float l = nl * INV_PI;
        return V(l, l, l);
    }
    const float a = 1.0f / (PI_F + sigma * (PI_F * 0.5f - 2.0f / 3.0f));
    const float b = sigma * a;
    const float nv = fmaxf(dot(n, in), 0.0f);
    float t = dot(out, in) - nl * nv;
    if (t > 0.0f) t = t / (fmaxf(nl, nv) + 1.17549435e-38f);
    const float single = a + b * t;
    const f3 albedo = vclamp(color, 0.0f, 1.0f);
    const float two_pi = 6.2831853071795864f;
    const float e_avg = a * PI_F + ((two_pi - 5.6f) / 3.0f) * b;
    const f3 e_ms = ((albedo * albedo) * INV_PI * (e_avg / (1.0f - e_avg))) /
                    (V(1.0f, 1.0f, 1.0f) - albedo * (1.0f - e_avg));
    const float e_in = a * PI_F + b * oren_nayar_g(nv);
    const f3 ms = e_ms * (1.0f - e_in);
    const float e_out = a * PI_F + b * oren_nayar_g(nl);
    return (V(single, single, single) + ms * (1.0f - e_out)) * nl;
}

struct Eval {
    f3 f;
    float pdf;
    uint32_t events;
};

__device__ Eval evaluate_mixture(const ClosureSet &cs, f3 point_incoming, f3 ns, f3 out_dir, uint32_t lobe_mask) {
    const f3 outgoing = safe_normalize(out_dir, ns);
    const f3 incoming = safe_normalize(point_incoming, neg(outgoing));
    f3 f = V(0.0f, 0.0f, 0.0f);
    float total = 0.0f, weighted = 0.0f;
    uint32_t events = 0u;
    const bool diffuse_enabled = (lobe_mask & EV_DIFFUSE) != 0u;
    const bool transparent_enabled = (lobe_mask & EV_TRANSPARENT) != 0u;
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i) {
        if (i >= cs.count) break;
        if (cs.kind[i] == KIND_DIFFUSE) {
            const float pdf = fmaxf(dot(cs.normal, outgoing), 0.0f) * INV_PI;
            const f3 value = (cs.weight[i] * diffuse_intensity(cs.normal, cs.roughness[i], cs.color[i], incoming,
                                                               outgoing)) * 1.0f;
            const f3 contrib = diffuse_enabled ? value : V(0.0f, 0.0f, 0.0f);
            const float w = diffuse_enabled ? cs.sw[i] : 0.0f;
            if (diffuse_enabled && fabsf(avg3(contrib)) > 0.0f) events |= EV_DIFFUSE | EV_REFLECTION;
            f = f + contrib;
            total += w;
            weighted += w * (diffuse_enabled ? pdf : 0.0f);
        } else {
            if (transparent_enabled) total += cs.sw[i];
        }
    }
    Eval e;
    e.f = f;
    e.pdf = total > 0.0f ? weighted / fmaxf(total, 1.0e-20f) : 0.0f;
    e.events = weighted > 0.0f ? events : 0u;
    return e;
}

__device__ __forceinline__ uint32_t runtime_flags(const ClosureSet &cs, bool back_facing) {
    uint32_t f = back_facing ? RT_BACKFACING : 0u;
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i) {
        if (i >= cs.count) break;
        f |= cs.kind[i] == KIND_DIFFUSE ? (RT_BSDF | RT_BSDF_HAS_EVAL) : (RT_BSDF | RT_TRANSPARENT);
    }
    return f;
}

// Shirley-Chiu concentric map (include/psycles/luisa/cycles_sample_mapping.h:49-62)
__device__ __forceinline__ void uniform_disk(float rx, float ry, float &dx, float &dy) {
    const float a = 2.0f * rx - 1.0f;
    const float b = 2.0f * ry - 1.0f;
    const bool x_major = a * a > b * b;
    const float safe_a = a != 0.0f ? a : 1.0f;
    const float safe_b = b != 0.0f ? b : 1.0f;
    const float radius = x_major ? a : b;
    const float angle = x_major ? 0.25f * PI_F * (b / safe_a) : 0.5f * PI_F - 0.25f * PI_F * (a / safe_b);
    dx = cosf(angle) * radius;
    dy = sinf(angle) * radius;
    if (a == 0.0f && b == 0.0f) dx = dy = 0.0f;
}

__device__ __forceinline__ f3 cosine_hemisphere(f3 n, float rx, float ry) {
    float dx, dy;
    uniform_disk(rx, ry, dx, dy);
    const float cosine = sqrtf(fmaxf(1.0f - (dx * dx + dy * dy), 0.0f));
    const f3 general = V(n.z - n.y, n.x - n.z, n.y - n.x);
    const f3 equal = V(n.z - n.y, n.x + n.z, -n.y - n.x);
    const bool use_general = (n.x != n.y) || (n.x != n.z);
    const f3 t = normalize(use_general ? general : equal);
    const f3 b = cross(n, t);
    return (t * dx + b * dy) + n * cosine;
}

struct BsdfSample {
    bool valid;
    f3 wi;
    Eval eval;
    uint32_t runtime_flags;
    uint32_t selected;
    float rescaled;  // the lobe random rescaled into the chosen closure (path traces)
};

__device__ BsdfSample sample_closures(const ClosureSet &cs, const Surface &sf, float u_lobe, float ux, float uy,
                                      uint32_t lobe_mask) {
    BsdfSample r;
    r.valid = false;
    r.selected = 0xffffffffu;
    r.rescaled = 0.0f;
    r.eval.f = V(0.0f, 0.0f, 0.0f);
    r.eval.pdf = 0.0f;
    r.eval.events = 0u;
    r.wi = V(0.0f, 0.0f, 0.0f);
    r.runtime_flags = runtime_flags(cs, sf.back_facing);
    float total = 0.0f;
    float weights[kMaxClosures];
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i) {
        weights[i] = 0.0f;
        if (i >= cs.count) continue;
        const bool eligible = cs.kind[i] == KIND_TRANSPARENT ? (lobe_mask & EV_TRANSPARENT) != 0u
                                                              : (lobe_mask & EV_DIFFUSE) != 0u;
        weights[i] = eligible ? cs.sw[i] : 0.0f;
        total += weights[i];
    }
    // categorical inversion (surface_closure_sampling_core.cpp:410-437)
    const float random = clampf(u_lobe, 0.0f, 0.99999994f);
    const float target = random * total;
    float accumulated = 0.0f;
    bool selected = false;
    uint32_t sel = 0u;
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i) {
        if (i >= cs.count) break;
        const float next = accumulated + weights[i];
        const bool choose = !selected && weights[i] > 0.0f && target < next;
        if (choose) {
            sel = i;
            r.rescaled = cs.count > 1u ? (target - accumulated) / fmaxf(weights[i], 1.0e-20f) : random;
        }
        selected |= choose;
        accumulated = next;
    }
    if (!selected) return r;
    r.selected = sel;
    uint32_t kind = KIND_DIFFUSE;
    f3 weight = V(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i) {
        if (i == sel) {
            kind = cs.kind[i];
            weight = cs.weight[i];
        }
    }
    const bool transparent = kind == KIND_TRANSPARENT;
    const f3 direction = transparent ? neg(sf.incoming) : cosine_hemisphere(cs.normal, ux, uy);
    const Eval mix = evaluate_mixture(cs, sf.incoming, sf.ns, direction, lobe_mask);
    const bool valid = transparent || dot(sf.ng, direction) > 0.0f;
    r.valid = valid;
    r.wi = direction;
    const f3 regular_f = valid ? mix.f : V(0.0f, 0.0f, 0.0f);
    const float regular_pdf = valid ? mix.pdf : 0.0f;
    float sel_weight = 0.0f;
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i)
        if (i == sel) sel_weight = weights[i];
    const f3 delta = (valid && transparent) ? weight * 1.0e6f : V(0.0f, 0.0f, 0.0f);
    const float delta_num = (valid && transparent) ? 1.0e6f * sel_weight : 0.0f;
    r.eval.f = regular_f + delta;
    r.eval.pdf = regular_pdf + delta_num / fmaxf(total, 1.0e-20f);
    const uint32_t events = transparent ? (EV_TRANSMISSION | EV_TRANSPARENT) : (EV_DIFFUSE | EV_REFLECTION);
    r.eval.events = valid ? events : 0u;
    return r;
}

__device__ void surface_aov(const ClosureSet &cs, f3 ns, f3 &albedo, f3 &normal, f3 &transparency) {
    f3 a = V(0.0f, 0.0f, 0.0f), n = V(0.0f, 0.0f, 0.0f), tr = V(0.0f, 0.0f, 0.0f);
    float total = 0.0f;
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i) {
        if (i >= cs.count) break;
        if (cs.kind[i] == KIND_DIFFUSE) {
            const float pw = fabsf(avg3(cs.weight[i]));
            a = a + cs.weight[i];
            n = n + cs.normal * pw;
            total += pw;
        } else {
            tr = tr + cs.weight[i];
        }
    }
    albedo = a;
    transparency = tr;
    normal = safe_normalize(total > 0.0f ? n : ns, ns);
}

// ===========================================================================
// triangle light sampling (include/psycles/luisa/spherical_geometry.h)
// ===========================================================================
__device__ __forceinline__ float triangle_solid_angle(f3 ref, f3 p0, f3 p1, f3 p2) {
    const f3 a = normalize_or(p0 - ref, V(1.0f, 0.0f, 0.0f));
    const f3 b = normalize_or(p1 - ref, V(0.0f, 1.0f, 0.0f));
    const f3 c = normalize_or(p2 - ref, V(0.0f, 0.0f, 1.0f));
    const float num = fabsf(dot(a, cross(b, c)));
    const float den = 1.0f + dot(a, b) + dot(b, c) + dot(c, a);
    return 2.0f * atan2f(num, den);
}

__device__ __forceinline__ bool use_solid_angle(f3 ref, f3 p0, f3 p1, f3 p2) {
    const f3 e0 = p1 - p0, e1 = p2 - p0, e2 = p2 - p1;
    const float longest = fmaxf(dot(e0, e0), fmaxf(dot(e1, e1), dot(e2, e2)));
    const f3 nn = cross(e0, e1);
    const float nl2 = dot(nn, nn);
    const float dist = dot(nn, p0 - ref) / fmaxf(nl2, 1.0e-20f);
    return (nl2 > 1.0e-20f) && (longest > dist * dist);
}

__device__ __forceinline__ bool triangle_pdf(f3 ref, f3 lp, f3 p0, f3 p1, f3 p2, float &value) {
    const f3 un = cross(p1 - p0, p2 - p0);
    const float da = sqrtf(fmaxf(dot(un, un), 0.0f));
    const float area = 0.5f * da;
    const f3 normal = un / fmaxf(da, 1.0e-20f);
    const f3 off = lp - ref;
    const float d2 = dot(off, off);
    const f3 dir = off / sqrtf(fmaxf(d2, 1.0e-20f));
    const float cosine = fabsf(dot(normal, neg(dir)));
    const bool solid = use_solid_angle(ref, p0, p1, p2);
    const float sa = triangle_solid_angle(ref, p0, p1, p2);
    const bool valid_area = (area > 0.0f) && (d2 > 1.0e-12f) && (cosine > 0.0f);
    value = solid ? 1.0f / fmaxf(sa, 1.0e-20f) : d2 / fmaxf(cosine * area, 1.0e-20f);
    return solid ? (valid_area && sa > 0.0f) : valid_area;
}

struct LightSample {
    f3 position, direction;
    float distance, pdf;
    bool valid;
};

__device__ LightSample sample_triangle(f3 ref, f3 p0, f3 p1, f3 p2, float rx, float ry) {
    const float x = clampf(rx, 0.0f, 1.0f), y = clampf(ry, 0.0f, 1.0f);
    const bool upper = y > x;
    const float bu = upper ? 0.5f * x : x - 0.5f * y;
    const float bv = upper ? y - 0.5f * x : 0.5f * y;
    const f3 apos = (p0 * (1.0f - bu - bv) + p1 * bu) + p2 * bv;
    const f3 aoff = apos - ref;
    const float ad2 = dot(aoff, aoff);
    const float ad = sqrtf(fmaxf(ad2, 1.0e-20f));
    const f3 adir = aoff / ad;
    LightSample r;
    const bool solid = use_solid_angle(ref, p0, p1, p2);
    if (solid) {
        const f3 a = normalize_or(p0 - ref, V(1.0f, 0.0f, 0.0f));
        const f3 b = normalize_or(p1 - ref, V(0.0f, 1.0f, 0.0f));
        const f3 c = normalize_or(p2 - ref, V(0.0f, 0.0f, 1.0f));
        const float cos_b = clampf(dot(a, c), -1.0f, 1.0f);
        const float cos_c = clampf(dot(a, b), -1.0f, 1.0f);
        const float sa = triangle_solid_angle(ref, p0, p1, p2);
        const f3 ab_n = normalize_or(cross(a, b), V(1.0f, 0.0f, 0.0f));
        const f3 ac_n = normalize_or(cross(a, c), V(0.0f, 1.0f, 0.0f));
        const float cos_alpha = clampf(dot(ab_n, ac_n), -1.0f, 1.0f);
        const float sin_alpha = sqrtf(fmaxf(1.0f - cos_alpha * cos_alpha, 0.0f));
        const float alpha = acosf(cos_alpha);
        const float phi = x * sa - alpha;
        const float sin_phi = sinf(phi);
        const float cos_phi = cosf(phi);
        const float u = cos_phi - cos_alpha;
        const float v = sin_phi + sin_alpha * cos_c;
        const float num = (v * cos_phi - u * sin_phi) * cos_alpha - v;
        const float den = (v * sin_phi + u * cos_phi) * sin_alpha;
        const float safe_den = fabsf(den) > 1.0e-20f ? den : (den >= 0.0f ? 1.0e-20f : -1.0e-20f);
        const float q = fabsf(den) <= 1.0e-20f ? 1.0f : num / safe_den;
        const f3 tangent = normalize_or(c - a * cos_b, V(0.0f, 1.0f, 0.0f));
        const float qc = clampf(q, -1.0f, 1.0f);
        const f3 c_prime = normalize_or(a * qc + tangent * sqrtf(fmaxf(1.0f - q * q, 0.0f)), a);
        const float z = 1.0f - y * (1.0f - dot(c_prime, b));
        const f3 edge = normalize_or(c_prime - b * dot(c_prime, b), a);
        const f3 sdir = normalize_or(b * z + edge * sqrtf(fmaxf(1.0f - z * z, 0.0f)), b);
        const f3 e01 = p1 - p0, e02 = p2 - p0;
        const f3 pv = cross(sdir, e02);
        const float det = dot(e01, pv);
        const float safe_det = fabsf(det) > 1.0e-20f ? det : (det >= 0.0f ? 1.0e-20f : -1.0e-20f);
        const float inv_det = 1.0f / safe_det;
        const f3 tv = ref - p0;
        const float su = dot(tv, pv) * inv_det;
        const f3 qv = cross(tv, e01);
        const float sv = dot(sdir, qv) * inv_det;
        const float sdist = dot(e02, qv) * inv_det;
        r.position = ref + sdir * sdist;
        r.direction = sdir;
        r.distance = sdist;
        const bool svalid = (fabsf(det) > 1.0e-20f) && (sdist > 0.0f) && (su >= -1.0e-5f) && (sv >= -1.0e-5f) &&
                            (su + sv <= 1.0f + 1.0e-5f) && (sa > 0.0f);
        float pdf;
        const bool pv_ok = triangle_pdf(ref, r.position, p0, p1, p2, pdf);
        r.pdf = pdf;
        r.valid = pv_ok && svalid;
    } else {
        r.position = apos;
        r.direction = adir;
        r.distance = ad;
        float pdf;
        const bool pv_ok = triangle_pdf(ref, r.position, p0, p1, p2, pdf);
        r.pdf = pdf;
        r.valid = pv_ok && ad2 > 1.0e-12f;
    }
    return r;
}

__device__ __forceinline__ bool sampled_side(uint32_t s, f3 ng, f3 dir) {
    const bool back = dot(ng, neg(dir)) < 0.0f;
    const bool front_ok = s == 1u || s == 2u || s == 4u;
    const bool back_ok = s == 1u || s == 3u || s == 4u;
    return back ? back_ok : front_ok;
}

// ===========================================================================
// light transport callables (src/luisa/path_tracer_lighting.cpp)
// ===========================================================================
__device__ __forceinline__ float power_heuristic(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 / fmaxf(a2 + b2, 1.0e-20f);
}
__device__ __forceinline__ float forward_light_weight(const KernelParams &k, float fpdf, float npdf, bool competing,
                                                      bool available) {
    if (k.mis_mode == 1u) return 1.0f;
    if (k.mis_mode == 2u) return (competing && available) ? 0.0f : 1.0f;
    return (competing && available) ? power_heuristic(fpdf, npdf) : 1.0f;
}
__device__ __forceinline__ float nee_light_weight(const KernelParams &k, float npdf, float fpdf) {
    if (k.mis_mode == 2u) return 1.0f;
    if (k.mis_mode == 1u) return fpdf <= 0.0f ? 1.0f : 0.0f;
    return power_heuristic(npdf, fpdf);
}
__device__ __forceinline__ f3 film_clamp(f3 c, uint32_t depth, float direct, float indirect) {
    c.x = isfinite(c.x) ? c.x : 0.0f;
    c.y = isfinite(c.y) ? c.y : 0.0f;
    c.z = isfinite(c.z) ? c.z : 0.0f;
    const float limit = depth > 0u ? indirect : direct;
    const float mag = fabsf(c.x) + fabsf(c.y) + fabsf(c.z);
    if (limit > 0.0f && mag > limit) return c * (limit / fmaxf(mag, 1.0e-20f));
    return c;
}
__device__ __forceinline__ float light_roulette(f3 u, float random, float inv) {
    const float maximum = fmaxf(fabsf(u.x), fmaxf(fabsf(u.y), fabsf(u.z)));
    const float probability = maximum * inv;
    const bool roulette = (inv > 0.0f) && (probability < 1.0f);
    const bool survives = !roulette || (random < probability);
    const float inverse = roulette ? 1.0f / fmaxf(probability, 1.0e-20f) : 1.0f;
    return survives ? inverse : 0.0f;
}

// ===========================================================================
// camera (src/luisa/path_tracer_camera.cpp:10-244, perspective)
// ===========================================================================
__device__ __forceinline__ float filter_lookup(const float *table, float random) {
    const float coordinate = clampf(random, 0.0f, 1.0f) * 1023.0f;
    uint32_t index = (uint32_t)coordinate;
    if (index > 1023u) index = 1023u;
    const uint32_t next = index + 1u < 1023u ? index + 1u : 1023u;
    const float t = coordinate - (float)index;
    const float a = table[index], b = table[next];
    return a + t * (b - a);
}

__device__ Ray camera_ray(const KernelParams &k, const float4 *sobol, const float *filter, uint32_t fx, uint32_t fy,
                          uint32_t sample, uint32_t &rng_hash) {
    const uint32_t cycles_y = k.full_height - 1u - fy;
    rng_hash = pixel_hash(fx, cycles_y, k.seed);
    float4 fs = sample4(sobol, k.sequence_size, sample, rng_hash, 0u);
    if (sample == 0u) {
        fs.x = 0.5f;
        fs.y = 0.5f;
    }
    const float jx = filter_lookup(filter, fs.x);
    const float jy = 1.0f - filter_lookup(filter, fs.y);
    const float width = (float)k.full_width, height = (float)k.full_height;
    const float sx = 2.0f * ((float)fx + jx) / width - 1.0f + 2.0f * k.shift_x;
    const float sy = 1.0f - 2.0f * ((float)fy + jy) / height + 2.0f * k.shift_y;
    f3 lo = V(0.0f, 0.0f, 0.0f), ld = V(0.0f, 0.0f, -1.0f);
    float clip_cos = 1.0f;
    if (k.projection == 1u) {
        // orthographic (src/luisa/path_tracer_camera.cpp:127-143)
        lo = V(sx * k.ortho_span * (width / height) * 0.5f, sy * k.ortho_span * 0.5f, 0.0f);
    } else {
        ld = normalize(V(sx * k.tan_h, sy * k.tan_v, -1.0f));
        clip_cos = -ld.z;
    }
    const f3 wo = cam_point(k.camera, lo);
    const f3 d = safe_normalize(cam_dir(k.camera, ld), V(0.0f, 0.0f, -1.0f));
    const float safe = fmaxf(fabsf(clip_cos), 1.0e-20f);
    const float c0 = k.near_clip / safe;
    const float c1 = (k.far_clip - k.near_clip) / safe;
    Ray r;
    r.o = wo + d * c0;
    r.d = d;
    r.tmin = 0.0f;
    r.tmax = c1;
    return r;
}
// ===========================================================================
// analytic lamps: sampling and forward intersection
// (include/psycles/luisa/analytic_light_sampling.h:84-1118,
//  include/psycles/luisa/analytic_light_intersection.h:23-260,
//  include/psycles/luisa/cycles_sample_mapping.h:331-387)
// ===========================================================================
constexpr float SM_PI = 3.14159265358979323846f;
constexpr float SM_INV_PI = 0.31830988618379067154f;
constexpr float SM_INV_TWO_PI = 0.15915494309189533577f;
constexpr float AL_PI = 3.1415926535897932f;
constexpr float AL_INV_PI = 0.3183098861837907f;
constexpr float RAY_MAXIMUM = 1.0e30f;

__device__ __forceinline__ f3 lamp3(const float *v) { return V(v[0], v[1], v[2]); }

// analytic_light_sampling::safe_normalize (:120-131)
__device__ __forceinline__ f3 lamp_normalize(f3 value, f3 fallback) {
    const float m2 = dot(value, value);
    return m2 > 0.0f ? value / sqrtf(fmaxf(m2, 1.0e-30f)) : fallback;
}

__device__ __forceinline__ float sin2_to_one_minus_cos(float s2) {
    return s2 > 0.0004f ? 1.0f - sqrtf(fmaxf(1.0f - s2, 0.0f)) : 0.5f * s2;
}

// make_orthonormals (include/psycles/luisa/cycles_sample_mapping.h)
__device__ __forceinline__ void make_orthonormals(f3 n, f3 &t, f3 &b) {
    const f3 general = V(n.z - n.y, n.x - n.z, n.y - n.x);
    const f3 equal = V(n.z - n.y, n.x + n.z, -n.y - n.x);
    const bool use_general = (n.x != n.y) || (n.x != n.z);
    t = normalize(use_general ? general : equal);
    b = cross(n, t);
}

struct Cone {
    f3 direction;
    float cosine, pdf;
};

// sample_uniform_cone (cycles_sample_mapping.h:356-378)
__device__ __forceinline__ Cone sample_uniform_cone(f3 axis, float omc, float rx, float ry) {
    float dx, dy;
    uniform_disk(rx, ry, dx, dy);
    const float r2 = dx * dx + dy * dy;
    const float cosine = 1.0f - r2 * omc;
    const float radial = sqrtf(fmaxf(omc * (2.0f - omc * r2), 0.0f));
    const float mx = dx * radial, my = dy * radial;
    f3 t, b;
    make_orthonormals(axis, t, b);
    const f3 fd = (t * mx + b * my) + axis * cosine;
    const bool finite = omc > 0.0f;
    Cone c;
    c.direction = finite ? fd : axis;
    c.cosine = finite ? cosine : 1.0f;
    c.pdf = finite ? SM_INV_TWO_PI / omc : 1.0f;
    return c;
}

__device__ __forceinline__ f3 sample_uniform_sphere(float rx, float ry) {
    const float z = 1.0f - 2.0f * rx;
    const float radius = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    const float phi = 2.0f * SM_PI * ry;
    return V(radius * cosf(phi), radius * sinf(phi), z);
}

struct PrimHit {
    bool valid;
    float distance;
    f3 position;
};

// intersect_sphere (:195-238), open interval
__device__ __forceinline__ PrimHit intersect_sphere(f3 ro, f3 rd, float tmin, float tmax, f3 c, float r) {
    const f3 co = c - ro;
    const float r2 = r * r;
    const float cd2 = dot(co, co);
    const float pd = dot(co, rd);
    const f3 perp = co - rd * pd;
    const float perp2 = dot(perp, perp);
    const float dist = pd - copysignf(sqrtf(fmaxf(r2 - perp2, 0.0f)), cd2 - r2);
    PrimHit h;
    h.valid = !((cd2 > r2) && (pd < 0.0f)) && (perp2 <= r2) && (dist > tmin) && (dist < tmax);
    h.distance = dist;
    h.position = ro + rd * dist;
    return h;
}

// intersect_disk (:240-273)
__device__ __forceinline__ PrimHit intersect_disk(f3 ro, f3 rd, float tmin, float tmax, f3 c, f3 n, float r) {
    const f3 co = ro - c;
    const float pdist = dot(co, n);
    const float cosine = dot(n, neg(rd));
    const float dist = cosine != 0.0f ? pdist / cosine : 0.0f;
    const f3 pos = ro + rd * dist;
    const f3 radial = pos - c;
    PrimHit h;
    h.valid = (pdist * cosine > 0.0f) && (dist >= 0.0f) && (dot(radial, radial) < r * r) && (dist > tmin) &&
              (dist < tmax);
    h.distance = dist;
    h.position = pos;
    return h;
}

// sphere_light_pdf (:281-307)
__device__ __forceinline__ float sphere_light_pdf(float cd2, float r2, f3 sn, f3 dir, bool had_transmission,
                                                  float outside_omc) {
    const float outside = SM_INV_TWO_PI / fmaxf(outside_omc, 1.0e-30f);
    float inside;
    if (had_transmission) {
        inside = 0.5f * SM_INV_TWO_PI;
    } else {
        const float c = dot(sn, dir);
        inside = c > 0.0f ? c * SM_INV_PI : 0.0f;
    }
    return cd2 > r2 ? outside : inside;
}

// spot_light_to_local (:309-320) through light_linear_transform (:137-166)
__device__ __forceinline__ f3 spot_light_to_local(const Lamp &l, f3 ray) {
    const f3 cx = lamp3(l.axis_x) * l.axis_scale[0];
    const f3 cy = lamp3(l.axis_y) * l.axis_scale[1];
    const f3 cz = lamp3(l.axis_z) * l.axis_scale[2];
    const f3 adj = cross(cy, cz);
    const float det = dot(cx, adj);
    const bool valid = fabsf(det) > 1.0e-20f;
    const float inv = 1.0f / (valid ? det : 1.0f);
    const f3 ix = adj * inv, iy = cross(cz, cx) * inv, iz = cross(cx, cy) * inv;
    f3 local = lamp_normalize(V(dot(ix, ray), dot(iy, ray), dot(iz, ray)), V(0.0f, 0.0f, -1.0f));
    local.z = -local.z;
    return local;
}

// spot_light_attenuation (:322-352)
__device__ __forceinline__ float spot_light_attenuation(f3 lr, float angle, float smooth) {
    const float ch = cosf(0.5f * angle);
    const float bw = (1.0f - ch) * smooth;
    const float lin = clampf(bw != 0.0f ? (lr.z - ch) / bw : 0.0f, 0.0f, 1.0f);
    const float sm = lin * lin * (3.0f - 2.0f * lin);
    const float hard = lr.z >= ch ? 1.0f : 0.0f;
    return bw > 0.0f ? sm : hard;
}

// spot_one_minus_cosine_larger_spread (:354-376)
__device__ __forceinline__ float spot_omc_larger_spread(const Lamp &l) {
    const float tn = tanf(0.5f * l.spot_angle);
    const float ts = fmaxf(l.axis_scale[0] * l.axis_scale[0], l.axis_scale[1] * l.axis_scale[1]);
    const float as = l.axis_scale[2] * l.axis_scale[2];
    const float c = 1.0f / sqrtf(1.0f + tn * tn * ts / fmaxf(as, 1.0e-30f));
    return 1.0f - c;
}

// point_disk_pdf (:856-873)
__device__ __forceinline__ float point_disk_pdf(float d2, float light_cosine, float r) {
    const float r2 = r * r;
    const float inverse_area = r2 > 0.0f ? 1.0f / fmaxf(r2 * AL_PI, 1.0e-20f) : 1.0f;
    return light_cosine > 0.0f ? inverse_area * d2 / light_cosine : 0.0f;
}

// point_eval_factor (:880-896)
__device__ __forceinline__ float point_eval_factor(float r, bool normalize) {
    const float r2 = r * r;
    const float finite_area = 4.0f * AL_PI * r2;
    const float area = r2 > 0.0f ? finite_area : 4.0f;
    const float inverse_area = normalize ? 1.0f / fmaxf(area, 1.0e-20f) : 1.0f;
    return inverse_area * AL_INV_PI;
}

struct LampSample {
    bool valid;
    f3 direction, position, normal;
    float distance, pdf, eval;
};

// sample_point_light (:898-979), sample_spot_light (:981-1105) and their
// sphere/disk geometry (:429-603).  SEGMENT: the spot proposal of a volume
// segment (VolumeAnalyticLightSampling::spot_from_segment,
// src/luisa/volume_analytic_light_sampling.cpp): outside a finite sphere it
// always samples the visible cap, never the spread cone, and it keeps a
// geometrically valid sample whose attenuation is zero.
template <bool SEGMENT = false>
__device__ LampSample sample_point_or_spot(const Lamp &l, f3 ref, f3 sn, bool has_transmission, float rx, float ry) {
    const bool spot = l.type == kLampSpot;
    const bool sphere = (l.flags & kLampSphere) != 0u;
    const f3 center = lamp3(l.position);
    const float r = l.radius;
    const f3 co = ref - center;
    const float cd2 = dot(co, co);
    const float r2 = r * r;
    const f3 light_normal = lamp_normalize(co, V(0.0f, 0.0f, -1.0f));
    const float sphere_cap = sin2_to_one_minus_cos(r2 / fmaxf(cd2, 1.0e-30f));
    f3 outside_axis = neg(light_normal);
    float outside_cap = sphere_cap;
    bool spread_cone = false;
    if (spot && !SEGMENT) {
        const float spot_cap = spot_omc_larger_spread(l);
        spread_cone = (cd2 > r2) && !(sphere_cap < spot_cap);
        if (spread_cone) {
            outside_axis = lamp3(l.axis_z);
            outside_cap = spot_cap;
        }
    }
    const bool finite_sphere = sphere && (r > 0.0f);
    LampSample s;
    if (finite_sphere) {
        // sample_sphere_geometry
        const float cd = sqrtf(fmaxf(cd2, 1.0e-30f));
        const f3 ln = co / cd;
        const bool outside = cd2 > r2;
        f3 dir = V(0.0f, 0.0f, 0.0f);
        float cosine = 1.0f, pdf = 0.0f, dist = 0.0f;
        f3 pos = ref;
        bool known = false;
        bool valid = (r > 0.0f) && (cd2 > 0.0f);
        if (outside) {
            const Cone cone = sample_uniform_cone(outside_axis, outside_cap, rx, ry);
            dir = cone.direction;
            cosine = cone.cosine;
            pdf = cone.pdf;
            if (spread_cone) {
                const PrimHit h = intersect_sphere(ref, dir, 0.0f, RAY_MAXIMUM, center, r);
                valid = valid && h.valid;
                dist = h.distance;
                pos = h.position;
                known = h.valid;
            }
        } else {
            if (has_transmission) {
                dir = sample_uniform_sphere(rx, ry);
                pdf = 0.5f * SM_INV_TWO_PI;
            } else {
                dir = cosine_hemisphere(sn, rx, ry);
                float dx, dy;
                uniform_disk(rx, ry, dx, dy);
                pdf = sqrtf(fmaxf(1.0f - (dx * dx + dy * dy), 0.0f)) * SM_INV_PI;
            }
            cosine = -dot(dir, ln);
        }
        if (!known) {
            dist = cd * cosine - copysignf(sqrtf(fmaxf(r2 - cd2 + cd2 * cosine * cosine, 0.0f)), cd2 - r2);
            pos = ref + dir * dist;
        }
        const f3 nrm = lamp_normalize(pos - center, neg(dir));
        s.normal = nrm;
        s.position = nrm * r + center;
        s.direction = dir;
        s.distance = dist;
        s.pdf = pdf;
        s.valid = valid;
    } else {
        // sample_disk_geometry
        const float cd = sqrtf(fmaxf(cd2, 1.0e-30f));
        const f3 ln = co / cd;
        f3 t, b;
        make_orthonormals(ln, t, b);
        float dx, dy;
        uniform_disk(rx, ry, dx, dy);
        const f3 pos = center + (t * dx + b * dy) * r;
        const f3 off = pos - ref;
        const float d2 = dot(off, off);
        const float d = sqrtf(fmaxf(d2, 1.0e-30f));
        const f3 dir = off / d;
        const float cosine = dot(ln, neg(dir));
        s.valid = (cd2 > 0.0f) && (cosine > 0.0f);
        s.direction = dir;
        s.position = pos;
        s.normal = neg(dir);
        s.distance = d;
        s.pdf = point_disk_pdf(d2, cosine, r);
    }
    s.eval = point_eval_factor(r, (l.flags & kLampNormalize) != 0u);
    if (spot) {
        const f3 local = spot_light_to_local(l, neg(s.direction));
        const float attenuation = spot_light_attenuation(local, l.spot_angle, l.spot_smooth);
        const bool use_attenuation = !finite_sphere || (cd2 > r2);
        s.eval = s.eval * (use_attenuation ? attenuation : 1.0f);
        if (!SEGMENT) s.valid = s.valid && (s.eval != 0.0f);
    }
    return s;
}

// sample_distant_light (:84-108)
__device__ __forceinline__ LampSample sample_distant_light(const Lamp &l, float rx, float ry) {
    const float half_angle = 0.5f * fmaxf(l.angle, 0.0f);
    const float sq = sinf(0.5f * half_angle);
    const float omc = 2.0f * sq * sq;
    const Cone ns = sample_uniform_cone(neg(lamp3(l.axis_z)), omc, rx, ry);
    const bool finite = half_angle > 0.0f;
    const float sh = sinf(half_angle);
    const float disk_area = AL_PI * sh * sh;
    LampSample s;
    s.direction = neg(ns.direction);
    s.pdf = ns.pdf;
    s.eval = ((l.flags & kLampNormalize) != 0u && finite) ? 1.0f / fmaxf(disk_area, 1.0e-20f) : 1.0f;
    s.position = neg(s.direction);
    s.normal = neg(s.direction);
    s.distance = RAY_MAXIMUM;
    s.valid = true;
    return s;
}

// ---- area lamps (src/luisa/area_light_sampling.cpp;
// include/psycles/luisa/analytic_light_sampling.h:605-846) ----------------
__device__ __forceinline__ float al_sdiv(float a, float b) { return b != 0.0f ? a / b : 0.0f; }

// area_spread_attenuation (:796-846)
__device__ float area_spread_attenuation(f3 dir_to_light, f3 light_normal, float spread) {
    const float cosine = fmaxf(dot(light_normal, neg(dir_to_light)), 0.0f);
    const float half_spread = 0.5f * fmaxf(spread, 0.0f);
    const float sine_angle = sqrtf(fmaxf(1.0f - cosine * cosine, 0.0f));
    const float tangent_angle = sine_angle / fmaxf(cosine, 1.0e-20f);
    const float tangent_spread = tanf(half_spread);
    const float normalization = half_spread > 0.05f
                                    ? 1.0f / fmaxf(tangent_spread - half_spread, 1.0e-20f)
                                    : 3.0f / fmaxf(half_spread * half_spread * half_spread, 1.0e-20f);
    const float finite_spread = fmaxf((tangent_spread - tangent_angle) * normalization, 0.0f);
    const float zero_spread = tangent_angle > 1.0e-5f ? 0.0f : AL_PI;
    const float narrowed = half_spread <= 0.0f ? zero_spread : finite_spread;
    return spread >= AL_PI - 1.0e-6f ? 1.0f : narrowed;
}

// rectangle_solid_angle_context (:605-690): the rectangle in the frame of
// its edges as seen from `ref` (Urena, Fajardo, King, "An area-preserving
// parametrization for spherical rectangles", EGSR 2013)
struct RectContext {
    f3 z;
    float z0, x0, x1, y0, y1;
    float e[4], a[4];
    float solid_angle, pdf;
};

__device__ RectContext rect_solid_angle_context(f3 ref, f3 center, f3 au, float lu, f3 av, float lv) {
    RectContext c;
    f3 z = cross(au, av);
    const f3 off = center - ref;
    float z0 = dot(off, z);
    if (z0 > 0.0f) {
        z = neg(z);
        z0 = -z0;
    }
    c.z = z;
    c.z0 = z0;
    const float cu = dot(off, au), cv = dot(off, av);
    c.x0 = cu - 0.5f * lu;
    c.x1 = cu + 0.5f * lu;
    c.y0 = cv - 0.5f * lv;
    c.y1 = cv + 0.5f * lv;
    const float raw[4] = {-c.y0, c.x1, c.y1, -c.x0};
#pragma unroll
    for (int i = 0; i < 4; ++i) c.e[i] = raw[i] / sqrtf(raw[i] * raw[i] + z0 * z0);
#pragma unroll
    for (int i = 0; i < 4; ++i) c.a[i] = asinf(fminf(fmaxf(-c.e[i] * c.e[(i + 1) & 3], -1.0f), 1.0f));
    c.solid_angle = -(c.a[0] + c.a[1] + c.a[2] + c.a[3]);
    const float min_e2 = fminf(fminf(c.e[0] * c.e[0], c.e[1] * c.e[1]), fminf(c.e[2] * c.e[2], c.e[3] * c.e[3]));
    const float dist = length(off);
    const float planar_pdf = al_sdiv(-(dist * dist * dist), z0 * lu * lv);
    const bool planar = (c.solid_angle < 1.0e-5f) || (min_e2 > 0.99999f);
    c.pdf = planar ? planar_pdf : 1.0f / c.solid_angle;
    return c;
}

// sample_rectangle_solid_angle (:716-794)
__device__ f3 sample_rect_solid_angle(f3 ref, f3 center, f3 au, float lu, f3 av, float lv, float rx, float ry,
                                      float &pdf) {
    const RectContext c = rect_solid_angle_context(ref, center, au, lu, av, lv);
    const float b0 = c.e[0], b1 = c.e[2];
    const float b0sq = b0 * b0;
    const float alpha = rx * c.solid_angle + c.a[2] + c.a[3];
    const float fu = al_sdiv(cosf(alpha) * b0 + b1, sinf(alpha));
    float cos_u = copysignf(1.0f / sqrtf(fu * fu + b0sq), fu);
    cos_u = fminf(fmaxf(cos_u, -1.0f), 1.0f);
    float su = -(cos_u * c.z0) / fmaxf(sqrtf(fmaxf(1.0f - cos_u * cos_u, 0.0f)), 1.0e-7f);
    su = fminf(fmaxf(su, c.x0), c.x1);
    const float du2 = su * su + c.z0 * c.z0;
    const float h0 = c.y0 / sqrtf(du2 + c.y0 * c.y0);
    const float h1 = c.y1 / sqrtf(du2 + c.y1 * c.y1);
    const float hv = h0 + ry * (h1 - h0);
    const float hv2 = hv * hv;
    const float sv = hv2 < 1.0f - 1.0e-6f ? hv * sqrtf(du2 / (1.0f - hv2)) : c.y1;
    pdf = c.pdf;
    return ((ref + au * su) + av * sv) + c.z * c.z0;
}

__device__ __forceinline__ float area_value(const Lamp &l) {
    float area = l.size_u * l.size_v;
    if ((l.flags & kLampEllipse) != 0u) area *= 0.25f * AL_PI;
    return area;
}

__device__ __forceinline__ float area_eval_factor(const Lamp &l, f3 direction, f3 normal) {
    const float area = area_value(l);
    const float inverse_area = ((l.flags & kLampNormalize) != 0u && area != 0.0f) ? al_sdiv(1.0f, area) : 1.0f;
    return AL_INV_PI * inverse_area * area_spread_attenuation(direction, normal, l.spread);
}

__device__ __forceinline__ float area_tangent_half_spread(float spread) {
    return spread == AL_PI ? FLT_MAX : tanf(0.5f * fmaxf(spread, 0.0f));
}

__device__ __forceinline__ float area_to_solid_angle(f3 normal, f3 direction, float distance_squared) {
    const float cosine = dot(normal, neg(direction));
    return cosine > 0.0f ? distance_squared / cosine : 0.0f;
}

// clamp_to_spread (area_light_sampling.cpp:179-420): the part of the lamp
// the spread cone about the reference point's foot can see, as a
// rectangle, a circle or the original shape
struct AreaShape {
    bool valid, rectangle;
    f3 center, au, av;
    float lu, lv;
};

__device__ AreaShape area_shape(const Lamp &l, f3 ref, f3 normal, float tangent) {
    AreaShape r;
    r.valid = true;
    r.rectangle = (l.flags & kLampEllipse) == 0u;
    r.center = lamp3(l.position);
    r.au = lamp3(l.axis_x);
    r.av = lamp3(l.axis_y);
    r.lu = l.size_u;
    r.lv = l.size_v;
    if ((l.flags & kLampFullSpread) != 0u) return r;
    const f3 center = r.center;
    const float plane_distance = dot(normal, ref - center);
    const f3 closest = ref - normal * plane_distance;
    const float spread_radius = plane_distance * tangent;
    const float spread_u = dot(r.au, closest - center);
    const float spread_v = dot(r.av, closest - center);
    const bool round = !r.rectangle && r.lu == r.lv;
    bool sample_spread = spread_radius == 0.0f;
    bool resolved = false;
    if (round && !sample_spread) {
        const float center_distance = sqrtf(spread_u * spread_u + spread_v * spread_v);
        const float light_radius = 0.5f * r.lu;
        if (center_distance >= light_radius + spread_radius) {
            r.valid = false;
        } else {
            sample_spread = (center_distance <= fabsf(light_radius - spread_radius)) && (spread_radius < light_radius);
            if (center_distance > fabsf(light_radius - spread_radius)) {
                const float rect_u = light_radius + spread_radius - center_distance;
                float rect_v;
                if (fabsf(light_radius * light_radius - spread_radius * spread_radius) >=
                    center_distance * center_distance) {
                    rect_v = 2.0f * fminf(light_radius, spread_radius);
                } else {
                    const float offset = center_distance +
                                         (spread_radius * spread_radius - light_radius * light_radius) /
                                             center_distance;
                    rect_v = sqrtf(4.0f * spread_radius * spread_radius - offset * offset);
                }
                const float rect_area = rect_u * rect_v;
                const float light_area = AL_PI * light_radius * light_radius;
                const float spread_area = AL_PI * spread_radius * spread_radius;
                if (rect_area < fminf(light_area, spread_area)) {
                    r.rectangle = true;
                    r.au = safe_normalize(center - closest, lamp3(l.axis_x));
                    r.av = cross(normal, r.au);
                    r.lu = rect_u;
                    r.lv = rect_v;
                    r.center = ((center + closest) + r.au * (spread_radius - light_radius)) * 0.5f;
                    resolved = true;
                } else {
                    sample_spread = spread_area < light_area;
                }
            }
        }
    } else if (!round && !sample_spread) {
        const float min_u = fmaxf(spread_u - spread_radius, -0.5f * r.lu);
        const float max_u = fminf(spread_u + spread_radius, 0.5f * r.lu);
        const float min_v = fmaxf(spread_v - spread_radius, -0.5f * r.lv);
        const float max_v = fminf(spread_v + spread_radius, 0.5f * r.lv);
        if (min_u >= max_u || min_v >= max_v) {
            r.valid = false;
        } else {
            const float rect_u = max_u - min_u, rect_v = max_v - min_v;
            const float rect_area = rect_u * rect_v;
            const float ellipse_area = r.rectangle ? FLT_MAX : 0.25f * AL_PI * r.lu * r.lv;
            const float spread_area = AL_PI * spread_radius * spread_radius;
            if (r.rectangle || rect_area < fminf(ellipse_area, spread_area)) {
                const float cu = 0.5f * (min_u + max_u), cv = 0.5f * (min_v + max_v);
                r.center = r.center + (r.au * cu + r.av * cv);
                r.lu = rect_u;
                r.lv = rect_v;
                r.rectangle = true;
                resolved = true;
            } else {
                r.rectangle = false;
                sample_spread = spread_area < ellipse_area;
            }
        }
    }
    if (r.valid && !resolved && sample_spread) {
        r.rectangle = false;
        r.center = center + (r.au * spread_u + r.av * spread_v);
        r.lu = 2.0f * spread_radius;
        r.lv = 2.0f * spread_radius;
    }
    return r;
}

// AreaLightSampling::from_position (:494-646): surface NEE
__device__ LampSample sample_area_light(const Lamp &l, f3 ref, float rx, float ry) {
    LampSample s;
    s.valid = false;
    s.direction = V(0.0f, 0.0f, 0.0f);
    s.position = lamp3(l.position);
    s.distance = 0.0f;
    s.pdf = 0.0f;
    s.eval = 0.0f;
    const f3 normal = neg(lamp3(l.axis_z));
    s.normal = normal;
    if (!(dot(s.position - ref, normal) <= 0.0f)) return s;
    const float tangent = area_tangent_half_spread(l.spread);
    const AreaShape sh = area_shape(l, ref, normal, tangent);
    if (!sh.valid) return s;
    f3 position = sh.center;
    float pdf = 0.0f;
    if (sh.rectangle) {
        position = sample_rect_solid_angle(ref, sh.center, sh.au, sh.lu, sh.av, sh.lv, rx, ry, pdf);
    } else if (tangent == 0.0f) {
        pdf = 1.0f;
    } else {
        float dx, dy;
        uniform_disk(rx, ry, dx, dy);
        position = sh.center + (sh.au * (0.5f * sh.lu * dx) + sh.av * (0.5f * sh.lv * dy));
        pdf = 4.0f * AL_INV_PI / (sh.lu * sh.lv);
    }
    const f3 off = position - ref;
    const float d2 = dot(off, off);
    const float d = sqrtf(fmaxf(d2, 0.0f));
    const f3 dir = off / fmaxf(d, 1.0e-30f);
    if (!sh.rectangle && tangent > 0.0f) pdf *= area_to_solid_angle(normal, dir, d2);
    const float factor = area_eval_factor(l, dir, normal);
    const f3 inplane = position - lamp3(l.position);
    const float u = dot(inplane, lamp3(l.axis_x)), v = dot(inplane, lamp3(l.axis_y));
    const float eu = (0.5f + 1.0e-7f) * l.size_u + 1.0e-6f;
    const float ev = (0.5f + 1.0e-7f) * l.size_v + 1.0e-6f;
    const bool inside_rect = fabsf(u) <= eu && fabsf(v) <= ev;
    const bool inside_ellipse = u * u / (eu * eu) + v * v / (ev * ev) <= 1.0f;
    s.valid = ((l.flags & kLampEllipse) != 0u ? inside_ellipse : inside_rect) && factor > 0.0f;
    s.direction = dir;
    s.position = position;
    s.distance = d;
    s.pdf = pdf;
    s.eval = factor;
    return s;
}

// intersect_area (analytic_light_intersection.h:271-353): the ray meets
// the lamp's front strictly inside (tmin, tmax), within its shape, with
// non-zero emission; `eval` from AreaLightSampling::from_intersection
__device__ bool area_candidate(const Lamp &l, f3 ro, f3 rd, float tmin, float tmax, float &distance, float &eval) {
    const f3 normal = neg(lamp3(l.axis_z));
    const f3 center = lamp3(l.position);
    const float dn = dot(rd, normal);
    const float t = al_sdiv(dot(center - ro, normal), dn);
    if (!(dn < 0.0f && t > tmin && t < tmax)) return false;
    const f3 inplane = (ro + rd * t) - center;
    const float u = al_sdiv(dot(inplane, lamp3(l.axis_x)), l.size_u);
    const float v = al_sdiv(dot(inplane, lamp3(l.axis_y)), l.size_v);
    const bool inside_rect = u >= -0.5f && u <= 0.5f && v >= -0.5f && v <= 0.5f;
    if (!inside_rect || ((l.flags & kLampEllipse) != 0u && !(u * u + v * v <= 0.25f))) return false;
    const AreaShape sh = area_shape(l, ro, normal, area_tangent_half_spread(l.spread));
    if (!sh.valid) return false;
    const float factor = area_eval_factor(l, rd, normal);
    if (!(area_value(l) > 0.0f && factor > 0.0f)) return false;
    distance = t;
    eval = factor;
    return true;
}

// AreaLightSampling::from_intersection (:648-726): the pdf of a forward hit
__device__ float area_forward_pdf(const Lamp &l, f3 ro, f3 rd, float distance) {
    const f3 normal = neg(lamp3(l.axis_z));
    const float tangent = area_tangent_half_spread(l.spread);
    const AreaShape sh = area_shape(l, ro, normal, tangent);
    if (!sh.valid) return 0.0f;
    if (sh.rectangle) return rect_solid_angle_context(ro, sh.center, sh.au, sh.lu, sh.av, sh.lv).pdf;
    if (tangent == 0.0f) return 1.0f;
    return 4.0f * AL_INV_PI / (sh.lu * sh.lv) * area_to_solid_angle(normal, rd, distance * distance);
}

// intersect_point / intersect_spot (analytic_light_intersection.h:33-260),
// split in two: lamp_candidate decides whether the ray hits the lamp nearer
// than `tmax` and with non-zero emission (the part every lamp of the forward
// scan needs); lamp_forward_pdf evaluates the sampling pdf for the one lamp
// the scan keeps.  Both compute the same expressions, in the same order, as
// the single-pass intersection.
__device__ __forceinline__ bool lamp_candidate(const Lamp &l, f3 ro, f3 rd, float tmin, float tmax, float &distance,
                                               float &eval) {
    if (l.type == kLampArea) return area_candidate(l, ro, rd, tmin, tmax, distance, eval);
    if (!(l.radius > 0.0f)) return false;
    const bool spot = l.type == kLampSpot;
    const bool sphere = (l.flags & kLampSphere) != 0u;
    const f3 center = lamp3(l.position);
    const f3 co = ro - center;
    PrimHit g;
    if (sphere) {
        g = intersect_sphere(ro, rd, tmin, tmax, center, l.radius);
    } else {
        const f3 light_normal = lamp_normalize(co, neg(rd));
        g = intersect_disk(ro, rd, tmin, tmax, center, light_normal, l.radius);
    }
    if (!g.valid) return false;
    distance = g.distance;
    eval = point_eval_factor(l.radius, (l.flags & kLampNormalize) != 0u);
    if (spot) {
        const float cd2 = dot(co, co);
        const float r2 = l.radius * l.radius;
        const f3 local = spot_light_to_local(l, neg(rd));
        const bool use_attenuation = !sphere || (cd2 > r2);
        const float attenuation = spot_light_attenuation(local, l.spot_angle, l.spot_smooth);
        eval = eval * (use_attenuation ? attenuation : 1.0f);
        return (dot(rd, co) < 0.0f) && (eval > 0.0f);
    }
    return eval > 0.0f;
}

__device__ __forceinline__ float lamp_forward_pdf(const Lamp &l, f3 ro, f3 rd, float distance, f3 previous_normal,
                                                  bool had_transmission) {
    if (l.type == kLampArea) return area_forward_pdf(l, ro, rd, distance);
    const bool spot = l.type == kLampSpot;
    const f3 center = lamp3(l.position);
    const f3 co = ro - center;
    const float cd2 = dot(co, co);
    const float r2 = l.radius * l.radius;
    if ((l.flags & kLampSphere) != 0u) {
        const float sphere_cap = sin2_to_one_minus_cos(r2 / fmaxf(cd2, 1.0e-30f));
        const float outside_cap = spot ? fminf(sphere_cap, spot_omc_larger_spread(l)) : sphere_cap;
        return sphere_light_pdf(cd2, r2, previous_normal, rd, had_transmission, outside_cap);
    }
    const f3 light_normal = lamp_normalize(co, neg(rd));
    const float light_cosine = dot(light_normal, neg(rd));
    return point_disk_pdf(distance * distance, light_cosine, l.radius);
}

// Forward-hit culling records of the first kLampCache lamps, staged in LDS
// once per block; lamps beyond them are read from HBM.
constexpr uint32_t kLampCache = 512u;

struct LampView {
    const float4 *sphere;  // centre, radius
    const uint32_t *bits;  // kLampBit* and the visibility mask
    uint32_t cached;       // records held by `sphere` / `bits`
};

// Whether a path step can meet a light: a lamp (forward hits, NEE) or an
// entry of the light distribution (NEE, emission MIS).
__host__ __device__ inline bool scene_has_lights(const SceneView &s) {
    return s.lamp_count > 0u || s.distribution_count > 0u || s.emitter_count > 0u;
}

// Every thread of the block must call this (it synchronises the block).
__device__ __forceinline__ LampView stage_lamps(const SceneView &s, float4 *lds_sphere, uint32_t *lds_bits) {
    const uint32_t n = s.lamp_count < kLampCache ? s.lamp_count : kLampCache;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        lds_sphere[i] = s.lamp_sphere[i];
        lds_bits[i] = s.lamp_bits[i];
    }
    __syncthreads();
    LampView v;
    v.sphere = lds_sphere;
    v.bits = lds_bits;
    v.cached = n;
    return v;
}

// Conservative reject for the forward scan: a sphere lamp and a disk lamp of
// radius r both lie in the ball of radius r about the lamp centre, so a ray
// that hits either in (tmin, tmax) passes within r of the centre at a ray
// parameter in [pd - r, pd + r].  The radius is widened by 1 % plus a margin
// proportional to the coordinate magnitudes, which covers the rounding of
// the exact intersection routines.
__device__ __forceinline__ bool lamp_may_hit(float4 sph, f3 ro, f3 rd, float tmin, float tmax) {
    const f3 co = V(sph.x - ro.x, sph.y - ro.y, sph.z - ro.z);
    const float pd = dot(co, rd);
    const f3 perp = co - rd * pd;
    const float perp2 = dot(perp, perp);
    const float scale = fabsf(ro.x) + fabsf(ro.y) + fabsf(ro.z) + fabsf(sph.x) + fabsf(sph.y) + fabsf(sph.z) +
                        fabsf(pd);
    const float r = sph.w * 1.01f + 2.0e-5f * scale;
    return (perp2 <= r * r) && (pd + r > tmin) && (pd - r < tmax);
}

// The lamp the forward-hit scan below would find among the lamps eligible
// for this ray (required bits, ray visibility, not `previous`), walking the
// lamp BVH instead of every lamp: the nearest candidate strictly before
// `best`, ties to the lowest lamp index (the scan keeps the first of equal
// distances, since each candidate must lie strictly before the current
// one).  Candidates are tested against one float beyond the current best so
// that an equal distance is seen and the tie rule applied.  Returns the lamp
// or 0xffffffff; `best` and `best_eval` follow the winner.
__device__ __forceinline__ uint32_t nearest_lamp_bvh(const SceneView &s, const LampView &lv, f3 ro, f3 rd, float tmin,
                                                     uint32_t required, uint32_t ray_visibility, uint32_t previous,
                                                     float &best, float &best_eval) {
    // zero direction components nudged so every slab distance is finite
    const f3 inv = V(1.0f / (fabsf(rd.x) > 1.0e-30f ? rd.x : copysignf(1.0e-30f, rd.x)),
                     1.0f / (fabsf(rd.y) > 1.0e-30f ? rd.y : copysignf(1.0e-30f, rd.y)),
                     1.0f / (fabsf(rd.z) > 1.0e-30f ? rd.z : copysignf(1.0e-30f, rd.z)));
    uint32_t found = 0xffffffffu;
    uint32_t node = 0u;
    while (node < s.lamp_node_count) {
        const LampNode n = s.lamp_nodes[node];
        const float limit = best == INFINITY ? best : __uint_as_float(__float_as_uint(best) + 1u);
        const float x0 = (n.lo[0] - ro.x) * inv.x, x1 = (n.hi[0] - ro.x) * inv.x;
        const float y0 = (n.lo[1] - ro.y) * inv.y, y1 = (n.hi[1] - ro.y) * inv.y;
        const float z0 = (n.lo[2] - ro.z) * inv.z, z1 = (n.hi[2] - ro.z) * inv.z;
        const float t_near = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
        const float t_far = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
        // loose on both ends: rounding of the slab distances never culls
        const float slack = 1.0e-4f * (fabsf(t_near) + fabsf(t_far)) + 1.0e-6f;
        const bool overlap = t_near <= t_far + slack && t_far + slack >= 0.0f && t_near - slack <= limit;
        if (!overlap) {
            node = n.miss;
            continue;
        }
        if (n.leaf == 0u) {
            node += 1u;
            continue;
        }
        const uint32_t first = n.leaf & 0xffffffu, count = n.leaf >> 24;
        for (uint32_t k = first; k < first + count; ++k) {
            const uint32_t li = s.lamp_order[k];
            const bool cached = li < lv.cached;
            const uint32_t bits = cached ? lv.bits[li] : s.lamp_bits[li];
            const bool eligible =
                (bits & required) == required && (bits & ray_visibility & 0xffu) != 0u && li != previous;
            if (!eligible) continue;
            const float lim = best == INFINITY ? best : __uint_as_float(__float_as_uint(best) + 1u);
            const float4 sph = cached ? lv.sphere[li] : s.lamp_sphere[li];
            if (!lamp_may_hit(sph, ro, rd, tmin, lim)) continue;
            float distance, eval;
            if (!lamp_candidate(s.lamps[li], ro, rd, tmin, lim, distance, eval)) continue;
            if (distance < best || (distance == best && found != 0xffffffffu && li < found)) {
                found = li;
                best = distance;
                best_eval = eval;
            }
        }
        node = n.miss;
    }
    return found;
}

// ===========================================================================
// light tree selection and reverse pdf
// (src/luisa/path_tracer_light_tree.cpp:14-826, surface variants)
// ===========================================================================
constexpr float LT_PI = 3.14159265358979323846f;
constexpr float LT_SAFE = 1.0e-20f;
constexpr uint32_t LT_INVALID = 0xffffffffu;

struct TreeImportance {
    float max, min;
};

__device__ __forceinline__ float lt_sin_from_cos(float c) { return sqrtf(fmaxf(1.0f - c * c, 0.0f)); }

__device__ __forceinline__ f3 lt_safe_normalize(f3 v, f3 fallback) {
    const float l2 = dot(v, v);
    return l2 > LT_SAFE ? v * (1.0f / sqrtf(fmaxf(l2, LT_SAFE))) : fallback;
}

// bounded_importance (:83-180)
__device__ TreeImportance lt_bounded(f3 n, bool has_transmission, f3 p2c, float cos_theta_u, f3 cone_axis,
                                     float theta_o, float theta_e, float4 trig, float distance, float energy) {
    cos_theta_u = clampf(cos_theta_u, -1.0f, 1.0f);
    const float sin_theta_u = lt_sin_from_cos(cos_theta_u);
    bool visible = energy > 0.0f;
    const float ic = has_transmission ? fabsf(dot(p2c, n)) : dot(p2c, n);
    const float is = lt_sin_from_cos(clampf(ic, -1.0f, 1.0f));
    const float minimum_incidence = ic >= cos_theta_u ? 1.0f : ic * cos_theta_u + is * sin_theta_u;
    const float maximum_incidence = fmaxf(ic * cos_theta_u - is * sin_theta_u, 0.0f);
    visible = visible && (has_transmission || minimum_incidence >= 0.0f);
    const float tc = clampf(dot(cone_axis, neg(p2c)), -1.0f, 1.0f);
    const float ts = lt_sin_from_cos(tc);
    const float tmu_c = tc * cos_theta_u + ts * sin_theta_u;
    const float to_c = trig.x, to_s = trig.y;
    const bool fully = (tc >= cos_theta_u) || (tmu_c >= to_c);
    const bool partially = !fully && ((theta_o + theta_e > LT_PI) || (tmu_c > trig.z));
    const float tmu_s = lt_sin_from_cos(clampf(tmu_c, -1.0f, 1.0f));
    const float minimum_outgoing = fully ? 1.0f : (partially ? tmu_c * to_c + tmu_s * to_s : 0.0f);
    visible = visible && (fully || partially);
    const float d = fmaxf(distance, LT_SAFE);
    const float geometric = 1.0f / (d * d);
    TreeImportance r;
    r.max = visible ? fabsf(minimum_incidence * energy * minimum_outgoing * geometric) : 0.0f;
    const float tpu_c = tc * cos_theta_u - ts * sin_theta_u;
    const bool minimum_is_zero = (theta_e - theta_o < 0.0f) || (tc < 0.0f) || (cos_theta_u < 0.0f) ||
                                 (tpu_c < trig.w);
    const float tpu_s = lt_sin_from_cos(clampf(tpu_c, -1.0f, 1.0f));
    const float maximum_outgoing = tpu_c * to_c - tpu_s * to_s;
    r.min = (minimum_is_zero || !visible) ? 0.0f : fabsf(maximum_incidence * energy * maximum_outgoing / (d * d));
    return r;
}

// measure_importance (:185-283)
__device__ TreeImportance lt_measure_importance(float4 bmin_energy, float4 bmax_theta_o, float4 axis_theta_e,
                                                float4 trig, uint32_t flags, f3 point, f3 n, bool has_transmission) {
    TreeImportance r;
    r.max = r.min = 0.0f;
    const float energy = bmin_energy.w;
    if (!((flags & kTreeHasOrientation) != 0u && energy > 0.0f)) return r;
    const f3 cone_axis = lt_safe_normalize(V(axis_theta_e.x, axis_theta_e.y, axis_theta_e.z), V(0.0f, 0.0f, 1.0f));
    const float theta_o = bmax_theta_o.w, theta_e = axis_theta_e.w;
    f3 p2c = neg(cone_axis);
    float cos_theta_u = trig.z;
    float distance = 1.0f;
    bool visible = true;
    if ((flags & kTreeDistant) == 0u) {
        const f3 bmin = V(bmin_energy.x, bmin_energy.y, bmin_energy.z);
        const f3 bmax = V(bmax_theta_o.x, bmax_theta_o.y, bmax_theta_o.z);
        const f3 centroid = (bmin + bmax) * 0.5f;
        const f3 extent = bmax - centroid;
        visible = has_transmission ||
                  (dot(n, centroid - point) +
                       dot(V(fabsf(n.x), fabsf(n.y), fabsf(n.z)), V(fabsf(extent.x), fabsf(extent.y), fabsf(extent.z))) >
                   0.0f);
        const f3 delta = centroid - point;
        distance = sqrtf(dot(delta, delta));
        p2c = lt_safe_normalize(delta, neg(cone_axis));
        // cos_bound_subtended_angle (:40-53)
        const f3 pc = point - centroid, bc = bmax - centroid;
        const float d2 = dot(pc, pc), r2 = dot(bc, bc);
        cos_theta_u = d2 <= r2 ? -1.0f : sqrtf(fmaxf(1.0f - r2 / fmaxf(d2, LT_SAFE), 0.0f));
        const f3 cb = centroid - bmax;
        distance = fmaxf(0.5f * sqrtf(dot(cb, cb)), distance);
    }
    if (visible)
        r = lt_bounded(n, has_transmission, p2c, cos_theta_u, cone_axis, theta_o, theta_e, trig, distance, energy);
    return r;
}

__device__ __forceinline__ TreeImportance lt_emitter_importance(const SceneView &s, uint32_t i, f3 point, f3 n,
                                                               bool has_transmission) {
    const TreeEmitter &e = s.tree_emitters[i];
    return lt_measure_importance(e.bounds_min_energy, e.bounds_max_theta_o, e.cone_axis_theta_e, e.trig, e.identity.y,
                                 point, n, has_transmission);
}

// child_importance (:317-350)
__device__ TreeImportance lt_child_importance(const SceneView &s, uint32_t index, f3 point, f3 n,
                                              bool has_transmission) {
    const TreeNode &node = s.tree_nodes[index];
    TreeImportance r;
    r.max = r.min = 0.0f;
    if (node.topology.w == kTreeInner || node.emitters.y > 1u)
        r = lt_measure_importance(node.bounds_min_energy, node.bounds_max_theta_o, node.cone_axis_theta_e, node.trig,
                                  node.emitters.z, point, n, has_transmission);
    else if (node.emitters.y == 1u)
        r = lt_emitter_importance(s, node.emitters.x, point, n, has_transmission);
    return r;
}

// left_probability (:352-386)
__device__ float lt_left_probability(const SceneView &s, uint32_t left, uint32_t right, f3 point, f3 n,
                                      bool has_transmission, bool &usable) {
    const TreeImportance li = lt_child_importance(s, left, point, n, has_transmission);
    const TreeImportance ri = lt_child_importance(s, right, point, n, has_transmission);
    const float max_total = li.max + ri.max;
    const float min_total = li.min + ri.min;
    const float max_p = li.max / fmaxf(max_total, LT_SAFE);
    const float min_p = min_total > 0.0f ? li.min / fmaxf(min_total, LT_SAFE)
                                         : 0.5f * ((li.max > 0.0f ? 1.0f : 0.0f) + (ri.max == 0.0f ? 1.0f : 0.0f));
    usable = max_total > 0.0f;
    return 0.5f * (max_p + min_p);
}

// select_leaf_emitter (:388-535) with reservoir_update (:388-416)
__device__ void lt_select_leaf(const SceneView &s, const TreeNode &node, f3 point, f3 n, bool has_transmission,
                               float &random, uint32_t &selected, float &pdf) {
    float sel_max = 0.0f, sel_min = 0.0f, tot_max = 0.0f, tot_min = 0.0f;
    uint32_t positive = 0u;
    const bool sample_maximum = random > 0.5f;
    const uint32_t first = node.emitters.x, count = node.emitters.y;
    if (count > 1u) random = random * 2.0f - (sample_maximum ? 1.0f : 0.0f);
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t idx = first + i;
        const TreeImportance imp = lt_emitter_importance(s, idx, point, n, has_transmission);
        tot_max += imp.max;
        tot_min += imp.min;
        positive += imp.max > 0.0f ? 1u : 0u;
        const float primary = sample_maximum ? imp.max : imp.min;
        const float primary_total = sample_maximum ? tot_max : tot_min;
        if (primary > 0.0f) {
            if (selected == LT_INVALID) {
                selected = idx;
                sel_max = imp.max;
                sel_min = imp.min;
            } else {
                const float threshold = primary / primary_total;
                if (random <= threshold) {
                    selected = idx;
                    sel_max = imp.max;
                    sel_min = imp.min;
                    random = random / threshold;
                } else {
                    random = (random - threshold) / fmaxf(1.0f - threshold, LT_SAFE);
                }
                random = clampf(random, 0.0f, 1.0f);
            }
        }
    }
    if (positive > 0u && tot_min == 0.0f) {
        if (sample_maximum) {
            sel_min = 1.0f;
            tot_min = (float)positive;
        } else {
            selected = LT_INVALID;
            float sel_uniform = 0.0f, tot_uniform = 0.0f;
            for (uint32_t i = 0; i < count; ++i) {
                const uint32_t idx = first + i;
                const TreeImportance imp = lt_emitter_importance(s, idx, point, n, has_transmission);
                const float weight = imp.max > 0.0f ? 1.0f : 0.0f;
                if (weight > 0.0f) {
                    tot_uniform += weight;
                    if (selected == LT_INVALID) {
                        selected = idx;
                        sel_uniform = weight;
                    } else {
                        const float threshold = weight / tot_uniform;
                        if (random <= threshold) {
                            selected = idx;
                            sel_uniform = weight;
                            random = random / threshold;
                        } else {
                            random = (random - threshold) / fmaxf(1.0f - threshold, LT_SAFE);
                        }
                        random = clampf(random, 0.0f, 1.0f);
                    }
                }
                if (selected == idx) {
                    sel_max = imp.max;
                    sel_min = 1.0f;
                }
            }
            tot_min = tot_uniform;
        }
    }
    if (selected != LT_INVALID)
        pdf *= 0.5f * (sel_max / fmaxf(tot_max, LT_SAFE) + sel_min / fmaxf(tot_min, LT_SAFE));
}

// make_sample_callable (:551-637): emitter id or LT_INVALID, and its pdf
__device__ uint32_t lt_sample(const SceneView &s, float u, f3 point, f3 n, bool has_transmission, float &pdf) {
    pdf = 0.0f;
    if (s.tree_node_count == 0u || s.tree_emitter_count == 0u || s.tree_root >= s.tree_node_count)
        return LT_INVALID;
    float random = clampf(u, 0.0f, 1.0f);
    float selection_pdf = 1.0f;
    uint32_t index = s.tree_root, selected = LT_INVALID, steps = 0u;
    bool active = true;
    while (active && steps <= s.tree_node_count) {
        steps += 1u;
        const TreeNode &node = s.tree_nodes[index];
        if (node.topology.w != kTreeInner) {
            lt_select_leaf(s, node, point, n, has_transmission, random, selected, selection_pdf);
            active = false;
        } else {
            bool usable;
            const float p_left =
                lt_left_probability(s, node.topology.y, node.topology.z, point, n, has_transmission, usable);
            if (!usable) {
                active = false;
            } else {
                const float p_right = 1.0f - p_left;
                const bool choose_right = random <= p_right;
                index = choose_right ? node.topology.z : node.topology.y;
                selection_pdf *= choose_right ? p_right : p_left;
                random = choose_right ? random / fmaxf(p_right, LT_SAFE) : (random - p_right) / fmaxf(p_left, LT_SAFE);
                random = clampf(random, 0.0f, 1.0f);
            }
        }
    }
    if (selected == LT_INVALID) return LT_INVALID;
    pdf = selection_pdf;
    return s.tree_emitters[selected].identity.x;
}

// make_pdf_callable (:639-744)
__device__ float lt_pdf(const SceneView &s, uint32_t emitter_id, f3 point, f3 n, bool has_transmission) {
    if (s.tree_node_count == 0u || s.tree_emitter_count == 0u || emitter_id >= s.tree_emitter_count) return 0.0f;
    const uint2 mapping = s.tree_mapping[emitter_id];
    const uint32_t target = mapping.x;
    uint32_t current = mapping.y;
    if (target >= s.tree_emitter_count || current >= s.tree_node_count) return 0.0f;
    const TreeNode &leaf = s.tree_nodes[current];
    float tmax = 0.0f, tmin = 0.0f, tot_max = 0.0f, tot_min = 0.0f;
    uint32_t positive = 0u;
    for (uint32_t i = 0; i < leaf.emitters.y; ++i) {
        const uint32_t idx = leaf.emitters.x + i;
        const TreeImportance imp = lt_emitter_importance(s, idx, point, n, has_transmission);
        tot_max += imp.max;
        tot_min += imp.min;
        positive += imp.max > 0.0f ? 1u : 0u;
        if (idx == target) {
            tmax = imp.max;
            tmin = imp.min;
        }
    }
    if (!(tmax > 0.0f)) return 0.0f;
    float result = 0.5f * (tmax / fmaxf(tot_max, LT_SAFE) +
                           (tot_min > 0.0f ? tmin / fmaxf(tot_min, LT_SAFE) : 1.0f / fmaxf((float)positive, 1.0f)));
    uint32_t steps = 0u;
    bool valid = true;
    while (valid && current != s.tree_root && steps <= s.tree_node_count) {
        steps += 1u;
        const uint32_t parent_index = s.tree_nodes[current].topology.x;
        if (parent_index >= s.tree_node_count) {
            valid = false;
            result = 0.0f;
        } else {
            const TreeNode &parent = s.tree_nodes[parent_index];
            bool usable;
            const float p_left =
                lt_left_probability(s, parent.topology.y, parent.topology.z, point, n, has_transmission, usable);
            if (!usable) {
                valid = false;
                result = 0.0f;
            } else {
                result *= current == parent.topology.y ? p_left : 1.0f - p_left;
                current = parent_index;
            }
        }
    }
    if (current != s.tree_root) result = 0.0f;
    return result;
}

// make_triangle_emitter_callable (:746-790)
__device__ uint32_t lt_triangle_emitter(const SceneView &s, uint32_t object, uint32_t cprim) {
    uint32_t first = 0u, length = s.emitter_count;
    while (length > 0u) {
        const uint32_t half = length >> 1u;
        const uint32_t middle = first + half;
        const uint32_t eo = s.emitters[middle].instance, ep = s.emitter_cprim[middle];
        if (eo < object || (eo == object && ep < cprim)) {
            first = middle + 1u;
            length = length - half - 1u;
        } else {
            length = half;
        }
    }
    if (first < s.emitter_count && s.emitters[first].instance == object && s.emitter_cprim[first] == cprim)
        return first;
    return LT_INVALID;
}

// ===========================================================================
// per-material tables evaluated once per parameter set
// ===========================================================================

// Transparent extinction of a material for shadow rays (the transparent
// closure weights of its tree, summed in closure order and clamped to [0, 1];
// src/luisa/path_tracer_geometry.cpp:75-174) and its constant emission.
__device__ __forceinline__ void material_tables(const SceneView &s, uint32_t material, float4 &shadow,
                                                float4 &emission) {
    ClosureSet cs;
    material_closures(s, material, V(0.0f, 0.0f, 1.0f), cs);
    f3 ext = V(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t i = 0; i < kMaxClosures; ++i)
        if (i < cs.count && cs.kind[i] == KIND_TRANSPARENT) ext = ext + cs.weight[i];
    ext = vclamp(ext, 0.0f, 1.0f);
    shadow = make_float4(ext.x, ext.y, ext.z, 0.0f);
    const f3 e = material_constant_emission(s, material);
    emission = make_float4(e.x, e.y, e.z, 0.0f);
}

// ===========================================================================
// the path integrator as a step function (src/luisa/path_kernel_pipeline.cpp:
// 73-185).  One step consumes the closest hit of the current ray, adds
// emission, applies Russian roulette, writes the primary AOVs, builds the NEE
// shadow ray and samples the continuation.  The megakernel runs the steps of
// one path back to back; the wavefront integrator keeps the same state in HBM
// between its trace and shade kernels.  Both add radiance in the same order.
// ===========================================================================
constexpr uint32_t PATH_PREV_DELTA = 1u;       // previous scatter was singular
constexpr uint32_t PATH_SHADOW_PENDING = 2u;   // a shadow ray awaits its result
constexpr uint32_t PATH_FINISHING = 4u;        // path ended, pending shadow only
constexpr uint32_t PATH_DONE = 8u;             // path complete

// Light-pass sums of one sample (src/luisa/path_kernel_setup.cpp:696-733)
// and the first surface's diffuse / glossy weights that split scattered
// light (src/luisa/path_kernel_surface_scatter.cpp:163-176).
struct LightSums {
    f3 dd, di, gd, gi, td, ti, emission, environment;
    f3 path_dw, path_gw;
};

struct PathVars {
    Ray ray;
    uint32_t src_obj, src_prim;
    f3 throughput;
    float prev_pdf;
    f3 mis_normal;  // shading normal of the last non-transparent scatter
    f3 radiance;
    float transparency;
    f3 normal, albedo;
    PathState ps;
    uint32_t sample, rng_hash, step, bits;
    LightSums lp;   // only read and written by the light-pass kernels
    float min_pdf;  // path traces only: smallest non-transparent BSDF pdf so far
};

struct ShadowRay {
    Ray ray;
    TraceQuery q;
    f3 L;            // throughput * surviving light estimate
    uint32_t depth;  // bounce at NEE time (clamp selection)
    uint32_t tdepth; // transparent bounces already taken
    bool active;
    f3 dw, gw;       // light passes: diffuse / glossy split of the contribution
    uint32_t event;  // path traces only: the surface event that sampled the light
};

// light-pass routing (src/luisa/path_tracer_lighting.cpp:107-191)
__device__ __forceinline__ f3 component_ratio(f3 n, f3 d) {
    return V(fabsf(d.x) > 1.0e-20f ? n.x / d.x : 0.0f, fabsf(d.y) > 1.0e-20f ? n.y / d.y : 0.0f,
             fabsf(d.z) > 1.0e-20f ? n.z / d.z : 0.0f);
}

// split_scattered_light: diffuse, glossy and the rest (transmission)
__device__ __forceinline__ void lp_scattered(LightSums &lp, f3 c, f3 dw, f3 gw, bool direct) {
    const f3 dc = c * dw, gc = c * gw;
    const f3 tc = (c - dc) - gc;
    if (direct) {
        lp.dd = lp.dd + dc;
        lp.gd = lp.gd + gc;
        lp.td = lp.td + tc;
    } else {
        lp.di = lp.di + dc;
        lp.gi = lp.gi + gc;
        lp.ti = lp.ti + tc;
    }
}

// emission or environment seen directly, otherwise scattered light
__device__ __forceinline__ void lp_route(LightSums &lp, bool environment, f3 c, uint32_t flag, uint32_t bounce) {
    if ((flag & FLAG_ANY_PASS) == 0u) {
        if (environment)
            lp.environment = lp.environment + c;
        else
            lp.emission = lp.emission + c;
    } else if ((flag & FLAG_SURFACE_PASS) != 0u) {
        lp_scattered(lp, c, lp.path_dw, lp.path_gw, bounce == 1u);
    }
}

// split_nee_light weights: at the first surface the BSDF's own component
// ratios, deeper the path weights.  Every scattering closure evaluated here
// is diffuse, so diffuse_f is the whole f and glossy_f is zero.
__device__ __forceinline__ void nee_split(const LightSums &lp, f3 f, uint32_t bounce, f3 &dw, f3 &gw) {
    const bool direct = bounce == 0u;
    dw = direct ? component_ratio(f, f) : lp.path_dw;
    gw = direct ? component_ratio(V(0.0f, 0.0f, 0.0f), f) : lp.path_gw;
}

__device__ __forceinline__ TraceQuery path_query(const PathVars &p) {
    TraceQuery q;
    q.visibility = contract_visibility(p.ps.visibility);
    q.src_obj = p.src_obj;
    q.src_prim = p.src_prim;
    q.light_obj = q.light_prim = 0xffffffffu;
    return q;
}

// camera ray and initial state (src/luisa/path_kernel_setup.cpp)
__device__ __forceinline__ void init_path(const KernelParams &k, const float4 *sobol, const float *filter,
                                          uint32_t fx, uint32_t fy, uint32_t sample, PathVars &p) {
    p.ray = camera_ray(k, sobol, filter, fx, fy, sample, p.rng_hash);
    p.src_obj = p.src_prim = 0xffffffffu;
    p.throughput = V(1.0f, 1.0f, 1.0f);
    p.prev_pdf = 0.0f;
    p.mis_normal = V(0.0f, 0.0f, 0.0f);
    p.radiance = V(0.0f, 0.0f, 0.0f);
    p.transparency = 0.0f;
    p.normal = V(0.0f, 0.0f, 0.0f);
    p.albedo = V(0.0f, 0.0f, 0.0f);
    p.ps.flag = FLAG_MIS_SKIP | FLAG_TRANSPARENT_BACKGROUND;
    p.ps.visibility = CVIS_CAMERA;
    p.ps.bounce = p.ps.diffuse_bounce = p.ps.glossy_bounce = p.ps.transmission_bounce = p.ps.transparent_bounce = 0u;
    p.ps.rng_offset = 16u;
    p.sample = sample;
    p.step = 0u;
    p.bits = PATH_PREV_DELTA | (k.max_path_steps == 0u ? PATH_DONE : 0u);
    p.min_pdf = FLT_MAX;
}

// transparent shadow continuation (src/luisa/path_tracer_geometry.cpp:75-174):
// given the closest hit of the shadow ray, either finish (returns true) or
// advance past a transparent surface
__device__ __forceinline__ bool shadow_hit(const SceneView &s, uint32_t transparent_max, const Hit &sh, Ray &sray,
                                           uint32_t &tdepth, f3 &transmittance, bool opaque = false) {
    if (!sh.valid) return true;
    if (opaque) {
        transmittance = V(0.0f, 0.0f, 0.0f);
        return true;
    }
    if (tdepth >= transparent_max) {
        transmittance = V(0.0f, 0.0f, 0.0f);
        return true;
    }
    const uint32_t material = hit_material(s, sh.inst, sh.prim);
    f3 ext = V(0.0f, 0.0f, 0.0f);
    if (material != kNone) {
        const float4 e = s.material_shadow[material];
        ext = V(e.x, e.y, e.z);
    }
    if (fmaxf(ext.x, fmaxf(ext.y, ext.z)) > 0.0f) {
        transmittance = transmittance * ext;
        tdepth += 1u;
        sray.tmin = next_float_up(sh.t);
        return false;
    }
    transmittance = V(0.0f, 0.0f, 0.0f);
    return true;
}

template <bool LP>
__device__ __forceinline__ void apply_shadow(const KernelParams &k, PathVars &p, f3 L, uint32_t depth,
                                             f3 transmittance, f3 dw, f3 gw) {
    if (transmittance.x > 0.0f || transmittance.y > 0.0f || transmittance.z > 0.0f) {
        const f3 c = film_clamp(L * transmittance, depth, k.clamp_direct, k.clamp_indirect);
        p.radiance = p.radiance + c;
        if (LP) lp_scattered(p.lp, c, dw, gw, depth == 0u);
    }
}

// Diagnostic path trace record, schema "psycles.cycles-path-trace" v2
// (include/psycles/luisa/path_trace_schema.h:10-121): 328 float4 slots per
// traced pixel and sample, 8 global slots then 72 per surface event for the
// first 4 events; w = 1 marks a slot that was written.
constexpr uint32_t kTraceSchemaVersion = 2u;
constexpr uint32_t kTraceGlobalSlots = 8u;
constexpr uint32_t kTraceEventSlots = 72u;
constexpr uint32_t kTraceMaxEvents = 4u;
constexpr uint32_t kTraceSlotCount = 328u;
constexpr uint32_t kTraceShadowBase = 296u, kTraceShadowSlots = 8u;

__device__ __forceinline__ void trace_slot(float4 *slots, uint32_t index, f3 v) {
    slots[index] = make_float4(v.x, v.y, v.z, 1.0f);
}

// trace_uint32: a 32-bit word as two exact 16-bit halves
__device__ __forceinline__ f3 trace_halves(uint32_t v, float z) {
    return V((float)(v & 0xffffu), (float)(v >> 16), z);
}

// surface-event slots of event `step`, or null past the traced events
__device__ __forceinline__ float4 *event_slots(float4 *slots, uint32_t step) {
    return slots && step < kTraceMaxEvents ? slots + kTraceGlobalSlots + step * kTraceEventSlots : nullptr;
}

// shadow-event slots of event `step`, or null past the traced events
__device__ __forceinline__ float4 *shadow_slots(float4 *slots, uint32_t step) {
    return slots && step < kTraceMaxEvents ? slots + kTraceShadowBase + step * kTraceShadowSlots : nullptr;
}

// an object or primitive identity, -1 when invalid
// (src/luisa/path_kernel_direct_light_trace.cpp:6-11)
__device__ __forceinline__ float trace_identity(uint32_t v) { return v != 0xffffffffu ? (float)v : -1.0f; }

// Cycles' exclusion flags of an object's ray visibility, not a shadow catcher
// (src/luisa/cycles_shader_identity.h:64-97)
__device__ __forceinline__ uint32_t object_visibility_flags(uint32_t v) {
    return ((v & VIS_CAMERA) ? 0u : (1u << 24)) | ((v & VIS_DIFFUSE) ? 0u : (1u << 27)) |
           ((v & VIS_GLOSSY) ? 0u : (1u << 26)) | ((v & VIS_TRANSMISSION) ? 0u : (1u << 25)) |
           ((v & VIS_VOLUME_SCATTER) ? 0u : (1u << 23)) | (1u << 22);
}

// Cycles shader identity of an emissive triangle: shader index, cast shadow,
// smooth normal, use MIS and the object's visibility exclusions
// (src/luisa/cycles_shader_identity.h:123-147)
__device__ __forceinline__ uint32_t emissive_shader_id(const SceneView &s, const Emitter &em) {
    const Instance &inst = s.instances[em.instance];
    const Mesh &mesh = s.meshes[inst.mesh];
    const bool smooth = mesh.smooth_offset != kNone && s.smooth[mesh.smooth_offset + em.prim] != 0u;
    return em.material | (1u << 30) | (smooth ? (1u << 31) : 0u) | (1u << 28) |
           object_visibility_flags(inst.visibility & 0x3fu);
}

// Cycles KernelLightType of a lamp (src/luisa/cycles_shader_identity.h:37-53)
__device__ __forceinline__ uint32_t cycles_light_type(uint32_t type) {
    return type == kLampPoint ? 0u : type == kLampDistant ? 1u : type == kLampArea ? 3u : 4u;
}

// Cycles ClosureType of closure c (src/luisa/graph_surface_scattering.cpp:373-420;
// include/psycles/luisa/cycles_closure.h:18-32)
__device__ __forceinline__ uint32_t cycles_closure_type(const ClosureSet &cs, uint32_t c) {
    if (cs.kind[c] == KIND_TRANSPARENT) return 30u;
    return cs.roughness[c] < 1.0e-5f ? 2u : 3u;
}

// global slots written after the camera ray (src/luisa/path_kernel_setup.cpp:540-560)
__device__ __forceinline__ void trace_globals(const KernelParams &k, const float4 *sobol, uint32_t fx, uint32_t fy,
                                              uint32_t sample, const PathVars &p, float4 *slots) {
    float4 fs = sample4(sobol, k.sequence_size, sample, p.rng_hash, 0u);
    if (sample == 0u) {
        fs.x = 0.5f;
        fs.y = 0.5f;
    }
    trace_slot(slots, 0u, V((float)kTraceSchemaVersion, (float)fx, (float)(k.full_height - 1u - fy)));
    trace_slot(slots, 1u, V((float)sample, (float)(p.rng_hash & 0xffffu), (float)(p.rng_hash >> 16)));
    trace_slot(slots, 2u, V(fs.x, fs.y, 0.0f));
    trace_slot(slots, 3u, V(0.0f, 0.0f, 0.0f));
    trace_slot(slots, 4u, p.ray.o);
    trace_slot(slots, 5u, p.ray.d);
    trace_slot(slots, 6u, V(p.ray.tmin, p.ray.tmax, 0.5f));
}

// The shadow ray of a NEE sample in the path trace: ray, range, source and
// light identities (record_shadow, src/luisa/path_kernel_direct_light_trace.cpp:152-208);
// the first hit and transmittance follow once it is traced.
__device__ __forceinline__ void trace_shadow_ray(float4 *slots, uint32_t step, ShadowRay &sr, bool skip_self,
                                                 bool cast_shadow) {
    sr.event = step;
    float4 *e = shadow_slots(slots, step);
    if (!e) return;
    trace_slot(e, 0u, sr.ray.o);
    trace_slot(e, 1u, sr.ray.d);
    trace_slot(e, 2u, V(sr.ray.tmin, sr.ray.tmax, cast_shadow ? 1.0f : 0.0f));
    trace_slot(e, 3u, V(trace_identity(sr.q.src_obj), trace_identity(sr.q.src_prim), skip_self ? 1.0f : 0.0f));
    trace_slot(e, 4u, V(trace_identity(sr.q.light_obj), trace_identity(sr.q.light_prim), 0.0f));
}

// One path step.  Returns true when the path continues with p.ray; sr.active
// is set when a shadow ray must be traced and applied before the next step.
// LIGHTS = false compiles a step for scenes with no lamp and an empty light
// distribution (scene_has_lights): no forward lamp scan, no NEE, emission
// without MIS, which is what the runtime tests select for such scenes.
template <bool TRACE, bool LP, bool LIGHTS = true>
__device__ bool path_step(const SceneView &s, const LampView &lv, const KernelParams &k, const float4 *sobol,
                          PathVars &p, const Hit &hit, ShadowRay &sr, uint32_t *ev, float4 *slots = nullptr) {
    sr.active = false;
    if (TRACE && ev) {
        ev[0] = 1u;
        ev[1] = p.ps.rng_offset;
        ev[2] = ev[3] = ev[4] = ev[5] = 0xffffffffu;
        ev[7] = 0u;
    }
    const float u_term = sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 0u).x;
    if (TRACE && ev) ev[6] = __float_as_uint(u_term);
    // Lamps are transparent closest events resolved before the mesh or
    // background event, without Sobol dimensions
    // (src/luisa/path_kernel_pipeline.cpp:79-137,
    //  src/luisa/path_kernel_closest_event.cpp:24-266,
    //  src/luisa/path_kernel_forward_light.cpp:59-117)
    if (LIGHTS && s.lamp_count > 0u) {
        const float surface_distance = hit.valid ? hit.t : p.ray.tmax;
        const uint32_t ray_visibility = contract_visibility(p.ps.visibility);
        const bool camera_path = (ray_visibility & VIS_CAMERA) != 0u;
        const bool had_transmission = (p.ps.flag & FLAG_MIS_HAD_TRANSMISSION) != 0u;
        uint32_t previous = 0xffffffffu;
        for (;;) {
            float event_distance = surface_distance;
            uint32_t found = 0xffffffffu;
            float best_eval = 0.0f;
            const uint32_t required = kLampBitIntersectable | (camera_path ? 0u : kLampBitMis);
            if (s.lamp_node_count > 0u)
                found = nearest_lamp_bvh(s, lv, p.ray.o, p.ray.d, p.ray.tmin, required, ray_visibility, previous,
                                         event_distance, best_eval);
            for (uint32_t li = 0; li < (s.lamp_node_count > 0u ? 0u : s.lamp_count); ++li) {
                const bool cached = li < lv.cached;
                const uint32_t bits = cached ? lv.bits[li] : s.lamp_bits[li];
                const bool eligible =
                    (bits & required) == required && (bits & ray_visibility & 0xffu) != 0u && li != previous;
                if (!eligible) continue;
                const float4 sph = cached ? lv.sphere[li] : s.lamp_sphere[li];
                if (!lamp_may_hit(sph, p.ray.o, p.ray.d, p.ray.tmin, event_distance)) continue;
                float distance, eval;
                if (lamp_candidate(s.lamps[li], p.ray.o, p.ray.d, p.ray.tmin, event_distance, distance, eval)) {
                    found = li;
                    best_eval = eval;
                    event_distance = distance;
                }
            }
            if (found == 0xffffffffu) break;
            const Lamp &l = s.lamps[found];
            const float lamp_pdf = lamp_forward_pdf(l, p.ray.o, p.ray.d, event_distance, p.mis_normal,
                                                    had_transmission);
            const float selection_pdf =
                k.use_tree ? lt_pdf(s, s.emitter_count + found, p.ray.o, p.mis_normal, had_transmission)
                           : s.light_selection_pdf;
            const float light_pdf = lamp_pdf * selection_pdf;
            const bool competing = p.ps.bounce > 0u && (p.bits & PATH_PREV_DELTA) == 0u;
            const float mis = forward_light_weight(k, p.prev_pdf, light_pdf, competing, light_pdf > 0.0f);
            const f3 light_radiance = lamp3(l.color) * (l.power * best_eval);
            const f3 lc = film_clamp((p.throughput * light_radiance) * mis, p.ps.bounce > 1u ? 1u : 0u,
                                     k.clamp_direct, k.clamp_indirect);
            p.radiance = p.radiance + lc;
            if (LP) lp_route(p.lp, false, lc, p.ps.flag, p.ps.bounce);
            p.ps.transparent_bounce += 1u;
            p.ray.tmin = next_float_up(event_distance);
            previous = found;
            if (p.ps.transparent_bounce >= k.transparent_max) return false;
        }
    }
    if (!hit.valid) {
        if (TRACE && ev) ev[0] |= 4u;
        // background (src/luisa/path_kernel_background_event.cpp:20-159)
        const f3 bg = V(s.background[0], s.background[1], s.background[2]);
        const f3 c = film_clamp((p.throughput * bg) * 1.0f, p.ps.bounce > 1u ? 1u : 0u, k.clamp_direct,
                                k.clamp_indirect);
        const bool world_visible = (s.world_visibility & contract_visibility(p.ps.visibility)) != 0u;
        const bool transparent_bg =
            k.transparent_background != 0u && (p.ps.flag & FLAG_TRANSPARENT_BACKGROUND) != 0u;
        if (transparent_bg) {
            p.transparency += (p.throughput.x + p.throughput.y + p.throughput.z) * (1.0f / 3.0f);
        } else if (world_visible) {
            p.radiance = p.radiance + c;
        }
        if (LP) lp_route(p.lp, true, world_visible ? c : V(0.0f, 0.0f, 0.0f), p.ps.flag, p.ps.bounce);
        return false;
    }
    if (TRACE && ev) {
        ev[0] |= 2u;
        ev[2] = hit.object;
        ev[3] = hit.cprim;
    }
    Surface sf;
    resolve_surface(s, hit, p.ray, sf);
    ClosureSet cs;
    material_closures(s, sf.material, sf.ns, cs);
    const bool terminate_after_transparent = (p.ps.flag & FLAG_TERMINATE_AFTER_TRANSPARENT) != 0u;
    const bool terminate_on_next_surface = (p.ps.flag & FLAG_TERMINATE_ON_NEXT_SURFACE) != 0u;
    const uint32_t lobe_mask = terminate_after_transparent ? EV_TRANSPARENT : 0x3fu;
    // surface emission and forward MIS (src/luisa/path_kernel_surface_shading.cpp:115-181)
    float emission_weight = 1.0f;
    if (LIGHTS && k.nee && s.emitter_count > 0u && sf.emission_sampling != 0u) {
        const bool competing = p.ps.bounce > 0u && (p.bits & PATH_PREV_DELTA) == 0u;
        const f3 oriented = sf.back_facing ? neg(sf.ng) : sf.ng;
        const float area = 0.5f * length(cross(sf.wp1 - sf.wp0, sf.wp2 - sf.wp0));
        float selection_pdf = area * s.triangle_area_pdf;
        if (k.use_tree)
            selection_pdf = lt_pdf(s, lt_triangle_emitter(s, sf.object, sf.cprim), p.ray.o, p.mis_normal,
                                   (p.ps.flag & FLAG_MIS_HAD_TRANSMISSION) != 0u);
        float tpdf;
        const bool tvalid = triangle_pdf(p.ray.o, sf.position, sf.wp0, sf.wp1, sf.wp2, tpdf);
        const f3 dir = normalize_or(sf.position - p.ray.o, V(0.0f, 0.0f, 1.0f));
        const bool side = sampled_side(sf.emission_sampling, oriented, dir);
        const float value = tpdf * selection_pdf;
        const bool valid = side && tvalid && area > 0.0f && value > 0.0f;
        const float light_pdf = valid ? value : 0.0f;
        emission_weight = forward_light_weight(k, p.prev_pdf, light_pdf, competing, light_pdf > 0.0f);
    }
    const f3 emitted = film_clamp((p.throughput * cs.emission) * emission_weight, p.ps.bounce > 1u ? 1u : 0u,
                                  k.clamp_direct, k.clamp_indirect);
    p.radiance = p.radiance + emitted;
    if (LP) lp_route(p.lp, false, emitted, p.ps.flag, p.ps.bounce);
    if (terminate_on_next_surface) return false;
    // Russian roulette (cycles_path_state.h:416-446)
    float continuation = 1.0f;
    {
        const bool transparent = (p.ps.flag & FLAG_TRANSPARENT) != 0u;
        const bool before_minimum =
            transparent ? p.ps.transparent_bounce <= k.transparent_min : p.ps.bounce <= k.min_bounces;
        const f3 t = p.throughput;
        const float roulette = fminf(sqrtf(fmaxf(fabsf(t.x), fmaxf(fabsf(t.y), fabsf(t.z)))), 1.0f);
        const float prob = before_minimum ? 1.0f : roulette;
        continuation = prob;
        if (TRACE && ev) ev[7] = __float_as_uint(prob);
        if (prob <= 0.0f) {
            if (TRACE && ev) ev[0] |= 8u;
            return false;
        }
        if (prob != 1.0f) {
            if (u_term >= prob) {
                if (TRACE && ev) ev[0] |= 8u;
                return false;
            }
            p.throughput = p.throughput / prob;
        }
        // surface event slots, written once the path survived the roulette
        // (src/luisa/path_kernel_surface_shading.cpp:270-340)
        if (TRACE && slots && p.step < kTraceMaxEvents) {
            float4 *e = slots + kTraceGlobalSlots + p.step * kTraceEventSlots;
            trace_slot(e, 0u, V((float)p.step, (float)p.ps.rng_offset, (float)p.ps.bounce));
            trace_slot(e, 1u, V((float)p.ps.transparent_bounce, (float)p.ps.diffuse_bounce,
                                (float)p.ps.glossy_bounce));
            trace_slot(e, 2u, V((float)p.ps.transmission_bounce, (float)(p.ps.visibility & 0xffffu),
                                (float)(p.ps.visibility >> 16)));
            trace_slot(e, 3u, trace_halves(p.ps.flag, prob));
            trace_slot(e, 4u, p.throughput);
            trace_slot(e, 5u, p.ray.o);
            trace_slot(e, 6u, p.ray.d);
            trace_slot(e, 7u, V(p.ray.tmin, p.ray.tmax, 0.5f));
            trace_slot(e, 8u, V(hit.t, hit.u, hit.v));
            trace_slot(e, 9u, V((float)hit.object, (float)hit.cprim, 1.0f));
            trace_slot(e, 10u, trace_halves(sf.material, (float)cs.count));
            trace_slot(e, 11u, sf.position);
            trace_slot(e, 12u, sf.ng);
            trace_slot(e, 13u, sf.ns);
            trace_slot(e, 38u, trace_halves(runtime_flags(cs, sf.back_facing), 0.0f));
            const float u_lightterm = sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 2u).x;
            trace_slot(e, 14u, V(u_term, k.light_inv_rr_threshold > 0.0f ? u_lightterm : 0.0f, 0.0f));
            const float4 ul = sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 1u);
            trace_slot(e, 15u, V(ul.x, ul.y, ul.z));
            const float4 ub = sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 3u);
            trace_slot(e, 16u, V(ub.x, ub.y, ub.z));
            // the shading point's closures (src/luisa/path_kernel_surface_shading.cpp:350-380)
            for (uint32_t c = 0; c < cs.count && c < kMaxClosures; ++c) {
                trace_slot(e, 48u + 3u * c, V((float)c, (float)cycles_closure_type(cs, c), cs.sw[c]));
                trace_slot(e, 49u + 3u * c, cs.weight[c]);
                trace_slot(e, 50u + 3u * c, cs.normal);
            }
        }
    }
    if (p.ps.bounce == 0u) {
        f3 aov_albedo, aov_normal, aov_tr;
        surface_aov(cs, sf.ns, aov_albedo, aov_normal, aov_tr);
        p.albedo = p.albedo + p.throughput * aov_albedo;
        const f3 alpha = vclamp(V(1.0f, 1.0f, 1.0f) - aov_tr, 0.0f, 1.0f);
        const float avg_alpha = (alpha.x + alpha.y + alpha.z) * (1.0f / 3.0f);
        const bool recorded = (p.ps.flag & FLAG_SINGLE_PASS_DONE) != 0u;
        const bool writes = !recorded && (k.pass_alpha_threshold == 0.0f || avg_alpha >= k.pass_alpha_threshold);
        if (writes) {
            p.normal = aov_normal;
            p.ps.flag |= FLAG_SINGLE_PASS_DONE;
        }
    }
    const uint32_t rflags = runtime_flags(cs, sf.back_facing);
    // NEE: emissive mesh component (src/luisa/path_kernel_nee_emissive_mesh.cpp:30-298)
    if (LIGHTS && k.nee && (rflags & RT_BSDF_HAS_EVAL) != 0u) {
        const float4 ul = sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 1u);
        uint32_t light;
        DistEntry entry;
        if (k.use_tree) {
            // light tree (src/luisa/path_kernel_pipeline.cpp:140-150)
            float tree_pdf;
            light = lt_sample(s, ul.z, sf.position, sf.ns, (rflags & RT_BSDF_HAS_TRANSMISSION) != 0u, tree_pdf);
            if (light != LT_INVALID && light < s.distribution_count) {
                entry = s.distribution[light];
                entry.selection_pdf = tree_pdf;
            } else {
                light = LT_INVALID;
                entry.kind = 3u;  // sentinel: nothing sampled
                entry.index = 0u;
                entry.selection_pdf = entry.cumulative = 0.0f;
            }
        } else {
            // flat-CDF light choice (src/luisa/path_tracer_light_distribution.cpp:12-52)
            uint32_t first = 0u, len = s.distribution_count + 1u;
            while (len > 0u) {
                const uint32_t half = len >> 1u;
                const uint32_t mid = first + half;
                if (ul.z < s.distribution[mid].cumulative) {
                    len = half;
                } else {
                    first = mid + 1u;
                    len = len - half - 1u;
                }
            }
            light = first > 0u ? first - 1u : 0u;
            if (light > s.distribution_count - 1u) light = s.distribution_count - 1u;
            entry = s.distribution[light];
        }
        if (entry.kind == 0u) {
            if (TRACE && ev) {
                ev[0] |= 16u;
                ev[4] = light;
            }
            const Emitter em = s.emitters[entry.index];
            const Instance &einst = s.instances[em.instance];
            const Mesh &emesh = s.meshes[einst.mesh];
            const uint32_t *et = s.indices + 3u * (emesh.index_offset + em.prim);
            const f3 lp0 = load3(s.positions + 3u * (emesh.vertex_offset + et[0]));
            const f3 lp1 = load3(s.positions + 3u * (emesh.vertex_offset + et[1]));
            const f3 lp2 = load3(s.positions + 3u * (emesh.vertex_offset + et[2]));
            const f3 p0 = mat_mul(einst.o2w, lp0, 1.0f);
            const f3 p1 = mat_mul(einst.o2w, lp1, 1.0f);
            const f3 p2 = mat_mul(einst.o2w, lp2, 1.0f);
            const f3 local_n = normalize_or(cross(lp1 - lp0, lp2 - lp0), V(0.0f, 0.0f, 1.0f));
            const f3 egn = normalize_or(mat_mul(einst.n2w, local_n, 0.0f), V(0.0f, 0.0f, 1.0f));
            const float earea = 0.5f * length(cross(p1 - p0, p2 - p0));
            const LightSample ls = sample_triangle(sf.position, p0, p1, p2, ul.x, ul.y);
            const bool side = sampled_side(em.emission_sampling, egn, ls.direction);
            const float lpdf = ls.pdf * entry.selection_pdf;
            float4 *te = TRACE ? event_slots(slots, p.step) : nullptr;
            if (TRACE && te) {
                // record_sample / record_failed_sample
                // (src/luisa/path_kernel_direct_light_trace.cpp:55-105)
                if (ls.valid && side && earea > 0.0f && lpdf > 0.0f) {
                    trace_slot(te, 17u, V(5.0f, (float)light, (float)(einst.prim_offset + em.prim)));
                    trace_slot(te, 18u, V((float)em.instance, 0.0f, 0.0f));
                    trace_slot(te, 40u, trace_halves(emissive_shader_id(s, em), 0.0f));
                    trace_slot(te, 19u, V(lpdf, entry.selection_pdf, 1.0f));
                    trace_slot(te, 20u, ls.direction);
                    trace_slot(te, 21u, ls.position);
                    trace_slot(te, 22u, egn);
                } else {
                    trace_slot(te, 17u, V(-1.0f, -1.0f, -1.0f));
                }
            }
            if (ls.valid && side && earea > 0.0f && lpdf > 0.0f) {
                const uint32_t eobj = em.instance;
                const uint32_t eprim = einst.prim_offset + em.prim;
                const bool is_transmission = dot(ls.direction, sf.ns) < 0.0f;
                const bool same = eprim == sf.cprim && eobj == sf.object;
                const f3 oriented = is_transmission ? neg(sf.ng) : sf.ng;
                const bool reject_self = same && dot(ls.direction, oriented) > 0.0f;
                if (!reject_self) {
                    const float4 le4 = s.material_emission[em.material];
                    const f3 le = V(le4.x, le4.y, le4.z);
                    const Eval evl = evaluate_mixture(cs, sf.incoming, sf.ns, ls.direction, lobe_mask);
                    const float mis = nee_light_weight(k, lpdf, evl.pdf);
                    const f3 weighted = (evl.f * le) * (mis / lpdf);
                    if (TRACE && te) {
                        trace_slot(te, 23u, V(ls.distance, evl.pdf, mis));
                        trace_slot(te, 41u, evl.f);
                        trace_slot(te, 42u, evl.f);
                        trace_slot(te, 43u, V(0.0f, 0.0f, 0.0f));
                        trace_slot(te, 44u, weighted);
                    }
                    if (any_nonzero(weighted)) {
                        const float u_lightterm =
                            sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 2u).x;
                        const float rr = light_roulette(weighted, u_lightterm, k.light_inv_rr_threshold);
                        const f3 surviving = weighted * rr;
                        if (TRACE && te) {
                            trace_slot(te, 45u, V(1.0f, 1.0f, 1.0f));
                            trace_slot(te, 46u, p.throughput * surviving);
                        }
                        if (any_nonzero(surviving)) {
                            const ShadowOrigin sso = surface_shadow_origin(s, sf, hit, ls.direction);
                            const f3 so = sso.position;
                            const f3 soff = ls.position - so;
                            const float sdist = sqrtf(fmaxf(dot(soff, soff), 1.0e-20f));
                            sr.ray.o = so;
                            sr.ray.d = soff / sdist;
                            sr.ray.tmin = 0.0f;
                            sr.ray.tmax = sdist;
                            sr.q.visibility = VIS_SHADOW;
                            sr.q.src_obj = sso.skip_self ? sf.object : 0xffffffffu;
                            sr.q.src_prim = sso.skip_self ? sf.cprim : 0xffffffffu;
                            sr.q.light_obj = eobj;
                            sr.q.light_prim = eprim;
                            sr.L = p.throughput * surviving;
                            sr.depth = p.ps.bounce;
                            sr.tdepth = p.ps.transparent_bounce;
                            sr.active = true;
                            if (LP) nee_split(p.lp, evl.f, p.ps.bounce, sr.dw, sr.gw);
                            if (TRACE) trace_shadow_ray(slots, p.step, sr, sso.skip_self, true);
                        }
                    }
                }
            }
        } else if (entry.kind == 1u) {
            // NEE: analytic lamp component (src/luisa/path_kernel_nee_analytic.cpp:98-410)
            if (TRACE && ev) {
                ev[0] |= 16u;
                ev[4] = light;
            }
            const Lamp &l = s.lamps[entry.index];
            const bool reached_max_bounces = p.ps.bounce > l.max_bounces;
            const bool distant = l.type == kLampDistant;
            const bool has_transmission = (rflags & RT_BSDF_HAS_TRANSMISSION) != 0u;
            const LampSample ls = distant                 ? sample_distant_light(l, ul.x, ul.y)
                                  : l.type == kLampArea ? sample_area_light(l, sf.position, ul.x, ul.y)
                                                        : sample_point_or_spot(l, sf.position, sf.ns,
                                                                               has_transmission, ul.x, ul.y);
            const f3 light_radiance = lamp3(l.color) * (l.power * ls.eval);
            const float light_pdf = ls.pdf * entry.selection_pdf;
            float4 *te = TRACE ? event_slots(slots, p.step) : nullptr;
            if (TRACE && te) {
                // record_sample / record_failed_sample
                // (src/luisa/path_kernel_nee_analytic.cpp:233-249, 413)
                if (!reached_max_bounces && ls.valid && light_pdf > 0.0f) {
                    trace_slot(te, 17u, V((float)cycles_light_type(l.type), (float)light, (float)entry.index));
                    trace_slot(te, 18u, V((float)0xffffffffu, 0.0f, 0.0f));
                    trace_slot(te, 40u, trace_halves(0xffffffffu, 0.0f));
                    trace_slot(te, 19u, V(light_pdf, entry.selection_pdf, ls.eval));
                    trace_slot(te, 20u, ls.direction);
                    trace_slot(te, 21u, ls.position);
                    trace_slot(te, 22u, ls.normal);
                } else {
                    trace_slot(te, 17u, V(-1.0f, -1.0f, -1.0f));
                }
            }
            if (!reached_max_bounces && ls.valid && light_pdf > 0.0f) {
                const Eval evl = evaluate_mixture(cs, sf.incoming, sf.ns, ls.direction, lobe_mask);
                const float cycles_mis = nee_light_weight(k, light_pdf, evl.pdf);
                const float mis = (l.flags & kLampForward) != 0u ? cycles_mis : 1.0f;
                const f3 weighted = (evl.f * light_radiance) * (mis / fmaxf(light_pdf, 1.0e-20f));
                if (TRACE && te) {
                    trace_slot(te, 23u, V(ls.distance, evl.pdf, cycles_mis));
                    trace_slot(te, 41u, evl.f);
                    trace_slot(te, 42u, evl.f);
                    trace_slot(te, 43u, V(0.0f, 0.0f, 0.0f));
                    trace_slot(te, 44u, weighted);
                }
                if (any_nonzero(weighted)) {
                    const float u_lightterm =
                        sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 2u).x;
                    const float rr = light_roulette(weighted, u_lightterm, k.light_inv_rr_threshold);
                    const f3 surviving = weighted * rr;
                    if (TRACE && te) {
                        trace_slot(te, 45u, V(1.0f, 1.0f, 1.0f));
                        trace_slot(te, 46u, p.throughput * surviving);
                    }
                    if (any_nonzero(surviving)) {
                        const ShadowOrigin sso = surface_shadow_origin(s, sf, hit, ls.direction);
                        const f3 so = sso.position;
                        const f3 foff = ls.position - so;
                        const float fdist = sqrtf(fmaxf(dot(foff, foff), 1.0e-20f));
                        sr.ray.o = so;
                        sr.ray.d = distant ? ls.direction : foff / fdist;
                        sr.ray.tmin = 0.0f;
                        sr.ray.tmax = distant ? RAY_MAXIMUM : fdist;
                        sr.q.visibility = VIS_SHADOW;
                        sr.q.src_obj = sso.skip_self ? sf.object : 0xffffffffu;
                        sr.q.src_prim = sso.skip_self ? sf.cprim : 0xffffffffu;
                        sr.q.light_obj = sr.q.light_prim = 0xffffffffu;
                        sr.L = p.throughput * surviving;
                        sr.depth = p.ps.bounce;
                        sr.tdepth = p.ps.transparent_bounce;
                        sr.active = true;
                        if (LP) nee_split(p.lp, evl.f, p.ps.bounce, sr.dw, sr.gw);
                        if (TRACE) trace_shadow_ray(slots, p.step, sr, sso.skip_self, (l.flags & kLampCastShadow) != 0u);
                    }
                }
            }
        }
    }
    // surface scatter (src/luisa/path_kernel_surface_scatter.cpp:16-324)
    const float4 ub = sample4(sobol, k.sequence_size, p.sample, p.rng_hash, p.ps.rng_offset + 3u);
    const BsdfSample bs = sample_closures(cs, sf, ub.z, ub.x, ub.y, lobe_mask);
    if (TRACE && ev) ev[5] = bs.selected;
    if (TRACE && bs.selected < cs.count) {
        // the sampled closure and BSDF sample (src/luisa/path_kernel_surface_scatter.cpp:95-130)
        float4 *te = event_slots(slots, p.step);
        if (te) {
            const uint32_t c = bs.selected;
            trace_slot(te, 24u, V((float)c, (float)cycles_closure_type(cs, c), cs.sw[c]));
            trace_slot(te, 25u, V(ub.x, ub.y, bs.rescaled));
            trace_slot(te, 26u, cs.weight[c]);
            trace_slot(te, 27u, cs.normal);
            trace_slot(te, 28u, V(bs.eval.pdf, bs.eval.pdf, (float)label_from_events(bs.eval.events)));
            trace_slot(te, 29u, bs.wi);
            trace_slot(te, 30u, bs.eval.f);
            const bool sampled = bs.valid;
            const float r = sampled && cs.kind[c] == KIND_DIFFUSE ? 1.0f : 0.0f;
            trace_slot(te, 31u, V(r, r, 1.0f));
        }
    }
    if (!bs.valid || bs.eval.pdf <= 0.0f) return false;
    if (TRACE && ev) ev[0] |= 32u;
    const bool transparent = (bs.eval.events & EV_TRANSPARENT) != 0u;
    const bool singular = (bs.eval.events & EV_SINGULAR) != 0u;
    if (TRACE && ev && transparent) ev[0] |= 64u;
    if (LP && p.ps.bounce == 0u && !transparent) {
        p.lp.path_dw = component_ratio(bs.eval.f, bs.eval.f);
        p.lp.path_gw = component_ratio(V(0.0f, 0.0f, 0.0f), bs.eval.f);
    }
    p.throughput = p.throughput * (bs.eval.f / bs.eval.pdf);
    const f3 t = p.throughput;
    if (any_nan(t) || t.x < 0.0f || t.y < 0.0f || t.z < 0.0f) return false;
    next_surface(p.ps, label_from_events(bs.eval.events), bs.runtime_flags, k);
    if (!transparent) {
        p.prev_pdf = bs.eval.pdf;
        if (TRACE) p.min_pdf = fminf(p.min_pdf, bs.eval.pdf);
        p.bits = singular ? (p.bits | PATH_PREV_DELTA) : (p.bits & ~PATH_PREV_DELTA);
        p.mis_normal = sf.ns;
        const f3 nd = normalize(bs.wi);
        p.ray.o = surface_ray_origin(s, sf, nd);
        p.ray.d = nd;
        p.ray.tmin = 0.0f;
        p.ray.tmax = 1.0e30f;
    } else {
        p.ray.tmin = next_float_up(hit.t);
    }
    p.src_obj = sf.object;
    p.src_prim = sf.cprim;
    // post-scatter slots (src/luisa/path_kernel_surface_scatter.cpp:293-320)
    if (TRACE && slots && p.step < kTraceMaxEvents) {
        float4 *e = slots + kTraceGlobalSlots + p.step * kTraceEventSlots;
        trace_slot(e, 32u, V((float)p.ps.bounce, (float)p.ps.transparent_bounce, (float)p.ps.rng_offset));
        trace_slot(e, 33u, p.throughput);
        trace_slot(e, 34u, p.ray.o);
        trace_slot(e, 35u, p.ray.d);
        trace_slot(e, 36u, trace_halves(p.ps.flag, 0.0f));
        trace_slot(e, 37u, V(p.prev_pdf, p.min_pdf, continuation));
        trace_slot(e, 39u, trace_halves(p.ps.visibility, 0.0f));
    }
    p.step += 1u;
    return p.step < k.max_path_steps;
}

}  // namespace
}  // namespace psx
