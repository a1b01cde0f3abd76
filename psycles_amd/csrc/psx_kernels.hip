// psx_kernels.hip - HIP kernels of the MI355X render path (gfx950).
//
// Device half of Psycles' render call: camera rays, two-level BVH4 traversal
// with an LDS stack and the exact Cycles closest-hit contract, closure-tree
// shading, emissive-triangle NEE with transparent shadows, Russian roulette
// and in-order film accumulation.  Each stage cites the reference file whose
// behaviour it reproduces.
//
// Floating point: compiled with -ffp-contract=off and correctly rounded
// division/sqrt; fmaf appears exactly where the reference fuses.  Helper
// conventions (bit patterns fixed by this implementation):
//   dot(a, b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(v) = v * (1 / sqrt(dot(v, v)))
//   M * (v, w)   = ((c0*x + c1*y) + c2*z) + c3*w
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

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
// traversal: two-level BVH4, LDS stack, exact Cycles closest-hit contract
// (src/luisa/path_kernel_scene_traversal.cpp:34-68, 212-493;
//  src/luisa/cycles_triangle_intersection_component.cpp:14-108)
// ===========================================================================
constexpr int kBlock = 128;
constexpr int kStackDepth = static_cast<int>(kTraversalStackDepth);

struct Ray {
    f3 o, d;
    float tmin, tmax;
};

struct Hit {
    uint32_t inst, prim, object, cprim;
    float t, u, v;
    bool valid;
};

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

constexpr uint32_t kTlasTag = 0x80000000u;
constexpr uint32_t kInstanceTag = 0x40000000u;

template <bool COUNT>
__device__ Hit closest_hit(const SceneView &s, const Ray &ray, const TraceQuery &q, uint32_t *stack,
                           uint32_t &node_count, uint32_t &tri_count) {
    Hit best;
    best.valid = false;
    best.t = ray.tmax;
    best.inst = best.prim = best.object = best.cprim = 0xffffffffu;
    best.u = best.v = 0.0f;
    const float lower = ray.tmin > 0.0f ? ray.tmin * 0.99999f : ray.tmin - 1.0e-30f;
    // world ray and its clamped reciprocal
    const f3 wd = V(clamp_bvh_dir(ray.d.x), clamp_bvh_dir(ray.d.y), clamp_bvh_dir(ray.d.z));
    f3 o = ray.o;
    f3 d = wd;
    f3 inv = V(1.0f / wd.x, 1.0f / wd.y, 1.0f / wd.z);
    bool in_blas = false;
    uint32_t inst_index = 0;
    uint32_t inst_vis = 0, inst_prim_offset = 0;
    int sp = 0;
    uint32_t current = s.has_tlas ? (s.tlas_root | kTlasTag) : kEmptySlot;
    const int lane = threadIdx.x;
    while (true) {
        if (current == kEmptySlot) {
            if (sp == 0) break;
            --sp;
            current = stack[sp * kBlock + lane];
        }
        const uint32_t entry = current;
        current = kEmptySlot;
        if ((entry & kTlasTag) != 0u && in_blas) {
            in_blas = false;
            o = ray.o;
            d = wd;
            inv = V(1.0f / wd.x, 1.0f / wd.y, 1.0f / wd.z);
        }
        if ((entry & kInstanceTag) != 0u) {
            // enter an instance: object-space ray with the Cycles FMA transform
            inst_index = entry & ~(kTlasTag | kInstanceTag);
            const Instance &inst = s.instances[inst_index];
            inst_vis = inst.visibility;
            if ((inst_vis & q.visibility) == 0u) continue;
            inst_prim_offset = inst.prim_offset;
            in_blas = true;
            if (inst.applied) {
                o = ray.o;
                d = wd;
            } else {
                o = affine_point(inst.w2o, ray.o);
                const f3 od = affine_dir(inst.w2o, ray.d);
                d = V(clamp_bvh_dir(od.x), clamp_bvh_dir(od.y), clamp_bvh_dir(od.z));
            }
            inv = V(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
            current = inst.blas_root;
            continue;
        }
        const bool tlas = (entry & kTlasTag) != 0u;
        const Node4 &n = s.nodes[entry & ~kTlasTag];
        if (COUNT) node_count++;
        const float upper = (best.valid ? best.t : ray.tmax) * 1.00001f + 1.0e-30f;
        float tn[4];
        bool hit[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float t0x = (n.lo_x[i] - o.x) * inv.x, t1x = (n.hi_x[i] - o.x) * inv.x;
            const float t0y = (n.lo_y[i] - o.y) * inv.y, t1y = (n.hi_y[i] - o.y) * inv.y;
            const float t0z = (n.lo_z[i] - o.z) * inv.z, t1z = (n.hi_z[i] - o.z) * inv.z;
            const float near = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
            const float far = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
            const float far_safe = far * 1.00001f + 1.0e-30f;
            tn[i] = near;
            hit[i] = n.count[i] != kEmptySlot && near <= far_safe && far_safe >= lower && near <= upper;
        }
        // leaves first: items are tested immediately
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!hit[i] || n.count[i] == 0u) continue;
            if (tlas) {
                // one instance per TLAS leaf item
                for (uint32_t k = 0; k < n.count[i]; ++k) {
                    const uint32_t instance = s.tlas_items[n.child[i] + k];
                    if (sp < kStackDepth) {
                        stack[sp * kBlock + lane] = instance | kTlasTag | kInstanceTag;
                        ++sp;
                    }
                }
                continue;
            }
            const uint32_t first = n.child[i];
            for (uint32_t k = 0; k < n.count[i]; ++k) {
                const LeafTri &tri = s.tris[first + k];
                if (COUNT) tri_count++;
                float t, bu, bv;
                const bool ok = pluecker(o, d, ray.tmin, ray.tmax, V(tri.p0x, tri.p0y, tri.p0z),
                                         V(tri.p1x, tri.p1y, tri.p1z), V(tri.p2x, tri.p2y, tri.p2z), t, bu, bv);
                if (!ok) continue;
                const uint32_t prim = __float_as_uint(tri.prim_bits);
                const uint32_t object = inst_index;
                const uint32_t cprim = inst_prim_offset + prim;
                const bool excluded = (object == q.src_obj && cprim == q.src_prim) ||
                                      (object == q.light_obj && cprim == q.light_prim);
                if (excluded || (inst_vis & q.visibility) == 0u) continue;
                if (!order_accepts(best, t, object, cprim)) continue;
                best.valid = true;
                best.t = t;
                best.u = bu;
                best.v = bv;
                best.inst = inst_index;
                best.prim = prim;
                best.object = object;
                best.cprim = cprim;
            }
        }
        // inner children: continue with the nearest, push the rest far-to-near
        uint32_t order[4];
        float keys[4];
        int m = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (hit[i] && n.count[i] == 0u) {
                order[m] = n.child[i] | (tlas ? kTlasTag : 0u);
                keys[m] = tn[i];
                ++m;
            }
        }
        // insertion sort, descending by entry distance
        for (int a = 1; a < m; ++a) {
            const uint32_t oa = order[a];
            const float ka = keys[a];
            int b = a - 1;
            while (b >= 0 && keys[b] < ka) {
                order[b + 1] = order[b];
                keys[b + 1] = keys[b];
                --b;
            }
            order[b + 1] = oa;
            keys[b + 1] = ka;
        }
        for (int a = 0; a + 1 < m; ++a) {
            if (sp < kStackDepth) {
                stack[sp * kBlock + lane] = order[a];
                ++sp;
            }
        }
        if (m > 0) current = order[m - 1];
    }
    return best;
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

__device__ __forceinline__ uint32_t material_of(const SceneView &s, const Instance &inst, const Mesh &mesh,
                                                uint32_t prim) {
    const uint32_t slot = mesh.tri_slot_offset != kNone ? s.tri_slots[mesh.tri_slot_offset + prim] : 0u;
    if (slot < inst.override_count) return s.overrides[inst.override_offset + slot];
    if (mesh.slot_count == 0u) return kNone;
    return s.slots[mesh.slot_offset + (slot < mesh.slot_count ? slot : mesh.slot_count - 1u)];
}

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
__device__ __forceinline__ f3 origin_with_self_exclusion(f3 wo, f3 wn, f3 oo, f3 od, f3 p0, f3 p1, f3 p2) {
    const f3 v0 = p0 - oo, v1 = p1 - oo, v2 = p2 - oo;
    const f3 e0 = v2 - v0, e1 = v0 - v1, e2 = v1 - v2;
    const float u = dot(cross(v2 + v0, e0), od);
    const float v = dot(cross(v0 + v1, e1), od);
    const float w = dot(cross(v1 + v2, e2), od);
    const float eps = FLT_EPSILON * fabsf(u + v + w);
    const float mn = fminf(u, fminf(v, w));
    const float mx = fmaxf(u, fmaxf(v, w));
    const bool covered = (mn >= eps) || (mx <= -eps);
    if (covered) return wo;
    return V(offset_component(wo.x, wn.x), offset_component(wo.y, wn.y), offset_component(wo.z, wn.z));
}

__device__ __forceinline__ f3 surface_ray_origin(const SceneView &s, const Surface &sf, f3 dir) {
    const Instance &inst = s.instances[sf.inst];
    const f3 co = sf.applied ? sf.position : sf.object_hit_position;
    const f3 cd = sf.applied ? dir : affine_dir(inst.w2o, dir);
    return origin_with_self_exclusion(sf.position, sf.ng, co, cd, sf.p0, sf.p1, sf.p2);
}

__device__ __forceinline__ f3 surface_shadow_origin(const SceneView &s, const Surface &sf, f3 dir) {
    const Instance &inst = s.instances[sf.inst];
    const f3 oo = sf.applied ? sf.position : affine_point(inst.w2o, sf.position);
    const f3 od = sf.applied ? dir : affine_dir(inst.w2o, dir);
    return origin_with_self_exclusion(sf.position, sf.ng, oo, od, sf.p0, sf.p1, sf.p2);
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

__device__ void material_closures(const SceneView &s, uint32_t material, f3 ns, ClosureSet &cs) {
    cs.count = 0u;
    cs.emission = V(0.0f, 0.0f, 0.0f);
    cs.normal = safe_normalize(ns, ns);
    if (material == kNone) return;
    const Material &m = s.materials[material];
    const float *p = s.params + m.param_offset;
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
        } else if (leaf.op == kLeafDiffuse) {
            f3 wt = vmax(color * w, V(0.0f, 0.0f, 0.0f));
            if (!(avg3(wt) >= 1.0e-5f)) wt = V(0.0f, 0.0f, 0.0f);
            const float allocation = fabsf(avg3(vmax(wt, V(0.0f, 0.0f, 0.0f))));
            if (allocation >= 1.0e-5f && cs.count < kMaxClosures) {
                const uint32_t c = cs.count++;
                cs.kind[c] = KIND_DIFFUSE;
                cs.weight[c] = vmax(wt, V(0.0f, 0.0f, 0.0f));
                cs.sw[c] = allocation;
                cs.roughness[c] = p[leaf.aux];
                cs.color[c] = color;
            }
        } else if (leaf.op == kLeafTransparent) {
            const f3 wt = color * w;
            const float sw = fabsf(avg3(wt));
            if (sw >= 1.0e-5f && cs.count < kMaxClosures) {
                const uint32_t c = cs.count++;
                cs.kind[c] = KIND_TRANSPARENT;
                cs.weight[c] = wt;
                cs.sw[c] = sw;
                cs.roughness[c] = 0.0f;
                cs.color[c] = V(1.0f, 1.0f, 1.0f);
            }
        }
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
};

__device__ BsdfSample sample_closures(const ClosureSet &cs, const Surface &sf, float u_lobe, float ux, float uy,
                                      uint32_t lobe_mask) {
    BsdfSample r;
    r.valid = false;
    r.selected = 0xffffffffu;
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
        if (choose) sel = i;
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
    const f3 ld = normalize(V(sx * k.tan_h, sy * k.tan_v, -1.0f));
    const float clip_cos = -ld.z;
    const f3 wo = cam_point(k.camera, V(0.0f, 0.0f, 0.0f));
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
// the path integrator (src/luisa/path_kernel_pipeline.cpp:73-185)
// ===========================================================================
struct PathResult {
    f3 radiance, normal, albedo;
    float transparency;
    uint32_t rays;
};

template <bool COUNT, bool TRACE>
__device__ PathResult integrate(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                                uint32_t fx, uint32_t fy, uint32_t sample, uint32_t *stack, uint32_t &nodes,
                                uint32_t &tris, uint32_t *trace) {
    uint32_t rng_hash;
    Ray ray = camera_ray(k, sobol, filter, fx, fy, sample, rng_hash);
    TraceQuery q;
    q.src_obj = q.src_prim = q.light_obj = q.light_prim = 0xffffffffu;
    q.visibility = VIS_CAMERA;
    f3 radiance = V(0.0f, 0.0f, 0.0f), throughput = V(1.0f, 1.0f, 1.0f);
    f3 sample_normal = V(0.0f, 0.0f, 0.0f), sample_albedo = V(0.0f, 0.0f, 0.0f);
    float transparency = 0.0f;
    bool primary_recorded = false;
    float previous_bsdf_pdf = 0.0f;
    bool previous_delta = true;
    PathState ps;
    ps.flag = FLAG_MIS_SKIP | FLAG_TRANSPARENT_BACKGROUND;
    ps.visibility = CVIS_CAMERA;
    ps.bounce = ps.diffuse_bounce = ps.glossy_bounce = ps.transmission_bounce = ps.transparent_bounce = 0u;
    ps.rng_offset = 16u;
    bool terminate_after_transparent = false, terminate_on_next_surface = false;
    uint32_t rays = 0u;
    for (uint32_t step = 0; step < k.max_path_steps; ++step) {
        uint32_t *ev = (TRACE && step < 8u) ? trace + step * 8u : nullptr;
        if (ev) {
            ev[0] = 1u;
            ev[1] = ps.rng_offset;
            ev[2] = ev[3] = ev[4] = ev[5] = 0xffffffffu;
            ev[7] = 0u;
        }
        const float u_term = sample4(sobol, k.sequence_size, sample, rng_hash, ps.rng_offset + 0u).x;
        const float4 ul = sample4(sobol, k.sequence_size, sample, rng_hash, ps.rng_offset + 1u);
        const float u_lightterm = sample4(sobol, k.sequence_size, sample, rng_hash, ps.rng_offset + 2u).x;
        const float4 ub = sample4(sobol, k.sequence_size, sample, rng_hash, ps.rng_offset + 3u);
        if (ev) ev[6] = __float_as_uint(u_term);
        // flat-CDF light choice (src/luisa/path_tracer_light_distribution.cpp:12-52)
        uint32_t light = 0xffffffffu;
        if (s.distribution_count > 0u) {
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
            uint32_t selected = first > 0u ? first - 1u : 0u;
            if (selected > s.distribution_count - 1u) selected = s.distribution_count - 1u;
            light = selected;
        }
        const Hit hit = closest_hit<COUNT>(s, ray, q, stack, nodes, tris);
        ++rays;
        if (!hit.valid) {
            if (ev) ev[0] |= 4u;
            // background (src/luisa/path_kernel_background_event.cpp:20-159)
            const f3 bg = V(s.background[0], s.background[1], s.background[2]);
            const f3 c = film_clamp((throughput * bg) * 1.0f, ps.bounce > 1u ? 1u : 0u, k.clamp_direct,
                                    k.clamp_indirect);
            const bool world_visible = (s.world_visibility & q.visibility) != 0u;
            const bool transparent_bg = k.transparent_background != 0u &&
                                        (ps.flag & FLAG_TRANSPARENT_BACKGROUND) != 0u;
            if (transparent_bg) {
                transparency += (throughput.x + throughput.y + throughput.z) * (1.0f / 3.0f);
            } else if (world_visible) {
                radiance = radiance + c;
            }
            break;
        }
        if (ev) {
            ev[0] |= 2u;
            ev[2] = hit.object;
            ev[3] = hit.cprim;
        }
        Surface sf;
        resolve_surface(s, hit, ray, sf);
        ClosureSet cs;
        material_closures(s, sf.material, sf.ns, cs);
        const uint32_t lobe_mask = terminate_after_transparent ? EV_TRANSPARENT : 0x3fu;
        // surface emission and forward MIS (src/luisa/path_kernel_surface_shading.cpp:115-181)
        float emission_weight = 1.0f;
        if (k.nee && s.emitter_count > 0u && sf.emission_sampling != 0u) {
            const bool competing = ps.bounce > 0u && !previous_delta;
            const f3 oriented = sf.back_facing ? neg(sf.ng) : sf.ng;
            const float area = 0.5f * length(cross(sf.wp1 - sf.wp0, sf.wp2 - sf.wp0));
            const float selection_pdf = area * s.triangle_area_pdf;
            float tpdf;
            const bool tvalid = triangle_pdf(ray.o, sf.position, sf.wp0, sf.wp1, sf.wp2, tpdf);
            const f3 dir = normalize_or(sf.position - ray.o, V(0.0f, 0.0f, 1.0f));
            const bool side = sampled_side(sf.emission_sampling, oriented, dir);
            const float value = tpdf * selection_pdf;
            const bool valid = side && tvalid && area > 0.0f && value > 0.0f;
            const float light_pdf = valid ? value : 0.0f;
            emission_weight = forward_light_weight(k, previous_bsdf_pdf, light_pdf, competing, light_pdf > 0.0f);
        }
        radiance = radiance + film_clamp((throughput * cs.emission) * emission_weight, ps.bounce > 1u ? 1u : 0u,
                                         k.clamp_direct, k.clamp_indirect);
        if (terminate_on_next_surface) break;
        // Russian roulette (cycles_path_state.h:416-446)
        {
            const bool transparent = (ps.flag & FLAG_TRANSPARENT) != 0u;
            const bool before_minimum = transparent ? ps.transparent_bounce <= k.transparent_min
                                                    : ps.bounce <= k.min_bounces;
            const float roulette = fminf(
                sqrtf(fmaxf(fabsf(throughput.x), fmaxf(fabsf(throughput.y), fabsf(throughput.z)))), 1.0f);
            const float p = before_minimum ? 1.0f : roulette;
            if (ev) ev[7] = __float_as_uint(p);
            if (p <= 0.0f) {
                if (ev) ev[0] |= 8u;
                break;
            }
            if (p != 1.0f) {
                if (u_term >= p) {
                    if (ev) ev[0] |= 8u;
                    break;
                }
                throughput = throughput / p;
            }
        }
        if (ps.bounce == 0u) {
            f3 aov_albedo, aov_normal, aov_tr;
            surface_aov(cs, sf.ns, aov_albedo, aov_normal, aov_tr);
            sample_albedo = sample_albedo + throughput * aov_albedo;
            const f3 alpha = vclamp(V(1.0f, 1.0f, 1.0f) - aov_tr, 0.0f, 1.0f);
            const float avg_alpha = (alpha.x + alpha.y + alpha.z) * (1.0f / 3.0f);
            const bool writes =
                !primary_recorded && (k.pass_alpha_threshold == 0.0f || avg_alpha >= k.pass_alpha_threshold);
            if (writes) {
                sample_normal = aov_normal;
                ps.flag |= FLAG_SINGLE_PASS_DONE;
            }
            primary_recorded |= writes;
        }
        const uint32_t rflags = runtime_flags(cs, sf.back_facing);
        // NEE: emissive mesh component (src/luisa/path_kernel_nee_emissive_mesh.cpp:30-298)
        if (k.nee && (rflags & RT_BSDF_HAS_EVAL) != 0u && light != 0xffffffffu &&
            s.distribution[light].kind == 0u) {
            if (ev) {
                ev[0] |= 16u;
                ev[4] = light;
            }
            const DistEntry entry = s.distribution[light];
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
            if (ls.valid && side && earea > 0.0f && lpdf > 0.0f) {
                const uint32_t eobj = em.instance;
                const uint32_t eprim = einst.prim_offset + em.prim;
                const bool is_transmission = dot(ls.direction, sf.ns) < 0.0f;
                const bool same = eprim == sf.cprim && eobj == sf.object;
                const f3 oriented = is_transmission ? neg(sf.ng) : sf.ng;
                const bool reject_self = same && dot(ls.direction, oriented) > 0.0f;
                if (!reject_self) {
                    const f3 le = material_constant_emission(s, em.material);
                    const Eval evl = evaluate_mixture(cs, sf.incoming, sf.ns, ls.direction, lobe_mask);
                    const float mis = nee_light_weight(k, lpdf, evl.pdf);
                    const f3 weighted = (evl.f * le) * (mis / lpdf);
                    if (any_nonzero(weighted)) {
                        const float rr = light_roulette(weighted, u_lightterm, k.light_inv_rr_threshold);
                        const f3 surviving = weighted * rr;
                        if (any_nonzero(surviving)) {
                            const f3 so = surface_shadow_origin(s, sf, ls.direction);
                            const f3 soff = ls.position - so;
                            const float sdist = sqrtf(fmaxf(dot(soff, soff), 1.0e-20f));
                            Ray sray;
                            sray.o = so;
                            sray.d = soff / sdist;
                            sray.tmin = 0.0f;
                            sray.tmax = sdist;
                            TraceQuery sq;
                            sq.visibility = VIS_SHADOW;
                            sq.src_obj = sf.object;
                            sq.src_prim = sf.cprim;
                            sq.light_obj = eobj;
                            sq.light_prim = eprim;
                            // trace_shadow (src/luisa/path_tracer_geometry.cpp:75-174)
                            f3 transmittance = V(1.0f, 1.0f, 1.0f);
                            uint32_t tdepth = ps.transparent_bounce;
                            for (;;) {
                                const Hit sh = closest_hit<COUNT>(s, sray, sq, stack, nodes, tris);
                                ++rays;
                                if (!sh.valid) break;
                                if (tdepth >= k.transparent_max) {
                                    transmittance = V(0.0f, 0.0f, 0.0f);
                                    break;
                                }
                                Surface ssf;
                                resolve_surface(s, sh, sray, ssf);
                                ClosureSet scs;
                                material_closures(s, ssf.material, ssf.ns, scs);
                                f3 ext = V(0.0f, 0.0f, 0.0f);
#pragma unroll
                                for (uint32_t i = 0; i < kMaxClosures; ++i)
                                    if (i < scs.count && scs.kind[i] == KIND_TRANSPARENT) ext = ext + scs.weight[i];
                                ext = vclamp(ext, 0.0f, 1.0f);
                                if (fmaxf(ext.x, fmaxf(ext.y, ext.z)) > 0.0f) {
                                    transmittance = transmittance * ext;
                                    tdepth += 1u;
                                    sray.tmin = next_float_up(sh.t);
                                } else {
                                    transmittance = V(0.0f, 0.0f, 0.0f);
                                    break;
                                }
                            }
                            if (transmittance.x > 0.0f || transmittance.y > 0.0f || transmittance.z > 0.0f) {
                                radiance = radiance + film_clamp((throughput * surviving) * transmittance, ps.bounce,
                                                                 k.clamp_direct, k.clamp_indirect);
                            }
                        }
                    }
                }
            }
        }
        // surface scatter (src/luisa/path_kernel_surface_scatter.cpp:16-324)
        const BsdfSample bs = sample_closures(cs, sf, ub.z, ub.x, ub.y, lobe_mask);
        if (ev) ev[5] = bs.selected;
        if (!bs.valid || bs.eval.pdf <= 0.0f) break;
        if (ev) ev[0] |= 32u;
        const bool transparent = (bs.eval.events & EV_TRANSPARENT) != 0u;
        const bool singular = (bs.eval.events & EV_SINGULAR) != 0u;
        if (ev && transparent) ev[0] |= 64u;
        throughput = throughput * (bs.eval.f / bs.eval.pdf);
        if (any_nan(throughput) || throughput.x < 0.0f || throughput.y < 0.0f || throughput.z < 0.0f) break;
        next_surface(ps, label_from_events(bs.eval.events), bs.runtime_flags, k);
        q.visibility = contract_visibility(ps.visibility);
        terminate_on_next_surface = (ps.flag & FLAG_TERMINATE_ON_NEXT_SURFACE) != 0u;
        terminate_after_transparent = (ps.flag & FLAG_TERMINATE_AFTER_TRANSPARENT) != 0u;
        if (!transparent) {
            previous_bsdf_pdf = bs.eval.pdf;
            previous_delta = singular;
            const f3 nd = normalize(bs.wi);
            ray.o = surface_ray_origin(s, sf, nd);
            ray.d = nd;
            ray.tmin = 0.0f;
            ray.tmax = 1.0e30f;
        } else {
            ray.tmin = next_float_up(hit.t);
        }
        q.src_obj = sf.object;
        q.src_prim = sf.cprim;
    }
    PathResult out;
    out.radiance = radiance;
    out.normal = sample_normal;
    out.albedo = sample_albedo;
    out.transparency = transparency;
    out.rays = rays;
    return out;
}

// wave-level reduction then one atomic per wave
__device__ __forceinline__ void add_counter(unsigned long long *dst, uint32_t value) {
    unsigned long long v = value;
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset, 64);
    if ((threadIdx.x & 63) == 0 && v != 0ull) atomicAdd(dst, v);
}

// 8x8 pixel tiles per wave: coherent primary rays share BVH nodes
__device__ __forceinline__ bool tile_pixel(const KernelParams &k, uint32_t &px, uint32_t &py) {
    const uint32_t wave = (blockIdx.x * (kBlock / 64)) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (k.window_width + 7u) >> 3;
    const uint32_t tx = wave % tiles_x, ty = wave / tiles_x;
    px = tx * 8u + (lane & 7u);
    py = ty * 8u + (lane >> 3);
    return px < k.window_width && py < k.window_height;
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void render_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                         const float *filter, FilmView film, uint32_t first,
                                                         uint32_t count, Counters *counters) {
    __shared__ uint32_t stack[kStackDepth * kBlock];
    uint32_t px, py;
    const bool active = tile_pixel(k, px, py);
    uint32_t rays = 0u, nodes = 0u, tris = 0u;
    if (active) {
        const size_t index = (size_t)py * k.window_width + px;
        float4 c = reinterpret_cast<float4 *>(film.combined)[index];
        float4 n = reinterpret_cast<float4 *>(film.normal)[index];
        float4 a = reinterpret_cast<float4 *>(film.albedo)[index];
        uint32_t completed = film.sample_count[index];
        for (uint32_t i = 0; i < count; ++i) {
            const PathResult r = integrate<COUNT, false>(s, k, sobol, filter, px + k.window_x, py + k.window_y,
                                                         first + i, stack, nodes, tris, nullptr);
            rays += r.rays;
            f3 rad = r.radiance;
            if (any_nan(rad)) rad = V(0.0f, 0.0f, 0.0f);
            c.x += rad.x;
            c.y += rad.y;
            c.z += rad.z;
            c.w += r.transparency;
            n.x += r.normal.x;
            n.y += r.normal.y;
            n.z += r.normal.z;
            n.w += 1.0f;
            a.x += r.albedo.x;
            a.y += r.albedo.y;
            a.z += r.albedo.z;
            a.w += 1.0f;
            completed += 1u;
        }
        reinterpret_cast<float4 *>(film.combined)[index] = c;
        reinterpret_cast<float4 *>(film.normal)[index] = n;
        reinterpret_cast<float4 *>(film.albedo)[index] = a;
        film.sample_count[index] = completed;
    }
    add_counter(&counters->rays, rays);
    if (COUNT) {
        add_counter(&counters->nodes, nodes);
        add_counter(&counters->tris, tris);
    }
}

__global__ __launch_bounds__(kBlock) void trace_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                        const float *filter, uint32_t sample, uint32_t *records) {
    __shared__ uint32_t stack[kStackDepth * kBlock];
    uint32_t px, py;
    if (!tile_pixel(k, px, py)) return;
    uint32_t *rec = records + ((size_t)py * k.window_width + px) * 64u;
    for (int i = 0; i < 64; ++i) rec[i] = 0u;
    uint32_t nodes = 0u, tris = 0u;
    integrate<false, true>(s, k, sobol, filter, px + k.window_x, py + k.window_y, sample, stack, nodes, tris, rec);
}

__global__ __launch_bounds__(kBlock) void primary_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                          const float *filter, uint32_t sample, uint32_t *records) {
    __shared__ uint32_t stack[kStackDepth * kBlock];
    uint32_t px, py;
    if (!tile_pixel(k, px, py)) return;
    uint32_t hash;
    const Ray ray = camera_ray(k, sobol, filter, px + k.window_x, py + k.window_y, sample, hash);
    TraceQuery q;
    q.visibility = VIS_CAMERA;
    q.src_obj = q.src_prim = q.light_obj = q.light_prim = 0xffffffffu;
    uint32_t nodes = 0u, tris = 0u;
    const Hit h = closest_hit<false>(s, ray, q, stack, nodes, tris);
    uint32_t *r = records + 4u * ((size_t)py * k.window_width + px);
    r[0] = h.valid ? h.object : 0xffffffffu;
    r[1] = h.valid ? h.cprim : 0xffffffffu;
    r[2] = h.valid ? __float_as_uint(h.t) : 0u;
    r[3] = h.valid ? 1u : 0u;
}

__global__ void sampler_kernel(KernelParams k, const float4 *sobol, uint32_t sample, uint32_t first_dim,
                               uint32_t dim_count, uint32_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.window_width * k.window_height) return;
    const uint32_t px = i % k.window_width, py = i / k.window_width;
    const uint32_t fx = px + k.window_x, fy = py + k.window_y;
    const uint32_t hash = pixel_hash(fx, k.full_height - 1u - fy, k.seed);
    uint32_t *r = out + (size_t)i * (1u + 4u * dim_count);
    r[0] = hash;
    for (uint32_t d = 0; d < dim_count; ++d) {
        const float4 v = sample4(sobol, k.sequence_size, sample, hash, first_dim + d);
        r[1 + 4 * d] = __float_as_uint(v.x);
        r[2 + 4 * d] = __float_as_uint(v.y);
        r[3 + 4 * d] = __float_as_uint(v.z);
        r[4 + 4 * d] = __float_as_uint(v.w);
    }
}

inline uint32_t tile_blocks(const KernelParams &k) {
    const uint32_t tiles = ((k.window_width + 7u) >> 3) * ((k.window_height + 7u) >> 3);
    const uint32_t waves_per_block = kBlock / 64;
    return (tiles + waves_per_block - 1u) / waves_per_block;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers called by the host library
// ---------------------------------------------------------------------------
hipError_t launch_render(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                         const FilmView &film, uint32_t first, uint32_t count, Counters *counters, bool count_nodes,
                         hipStream_t stream) {
    const uint32_t blocks = tile_blocks(k);
    if (blocks == 0u) return hipSuccess;
    if (count_nodes)
        hipLaunchKernelGGL(render_kernel<true>, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, film,
                           first, count, counters);
    else
        hipLaunchKernelGGL(render_kernel<false>, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, film,
                           first, count, counters);
    return hipGetLastError();
}

hipError_t launch_trace(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                        uint32_t sample, uint32_t *records, hipStream_t stream) {
    const uint32_t blocks = tile_blocks(k);
    if (blocks == 0u) return hipSuccess;
    hipLaunchKernelGGL(trace_kernel, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, sample, records);
    return hipGetLastError();
}

hipError_t launch_primary(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                          uint32_t sample, uint32_t *records, hipStream_t stream) {
    const uint32_t blocks = tile_blocks(k);
    if (blocks == 0u) return hipSuccess;
    hipLaunchKernelGGL(primary_kernel, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, sample, records);
    return hipGetLastError();
}

hipError_t launch_sampler(const KernelParams &k, const float4 *sobol, uint32_t sample, uint32_t first_dim,
                          uint32_t dim_count, uint32_t *out, hipStream_t stream) {
    const uint32_t n = k.window_width * k.window_height;
    if (n == 0u) return hipSuccess;
    hipLaunchKernelGGL(sampler_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, k, sobol, sample, first_dim,
                       dim_count, out);
    return hipGetLastError();
}

}  // namespace psx
