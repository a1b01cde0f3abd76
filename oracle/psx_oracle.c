/*
 * psx_oracle.c - CPU restatement of Psycles' device render path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP render path, never part
 * of the product.  Plain C99 compiled with -ffp-contract=off and without
 * fast-math: every float expression below is evaluated in the order written,
 * and fmaf() appears exactly where the reference fuses.  Transcendentals
 * (sinf, cosf, acosf, atan2f, tanf, expf) come from the host libm, as they do
 * in the reference's LLVM fallback backend.
 *
 * Conventions shared with the reference's Luisa lowering (the bit patterns of
 * these helpers are not pinned by any reference vector; they are the stated
 * contract of this restatement):
 *   dot(a, b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross(a, b)    = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   normalize(v)   = v * (1 / sqrt(dot(v, v)))
 *   M * (v, w)     = ((c0*x + c1*y) + c2*z) + c3*w   (column-major float4x4)
 *   lerp(a, b, t)  = a + t * (b - a)
 *   min/max        = fminf/fmaxf
 */
#define _GNU_SOURCE
#include "psx_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ======================================================================= */
/* small vector algebra                                                     */
/* ======================================================================= */

typedef struct { float x, y, z; } v3;
typedef struct { float x, y; } v2;

static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdiv(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 vscale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
              a.x * b.y - a.y * b.x);
}
static inline float vlength(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vnormalize(v3 a) {
    const float inv = 1.0f / sqrtf(vdot(a, a));
    return vscale(a, inv);
}
static inline v3 vsel(v3 a, v3 b, int c) { return c ? b : a; }
static inline float fsel(float a, float b, int c) { return c ? b : a; }
static inline float fclamp(float x, float lo, float hi) {
    return fminf(fmaxf(x, lo), hi);
}
static inline v3 vmax(v3 a, v3 b) {
    return V3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z));
}
static inline v3 vclamp(v3 a, float lo, float hi) {
    return V3(fclamp(a.x, lo, hi), fclamp(a.y, lo, hi), fclamp(a.z, lo, hi));
}
static inline int vany_nonzero(v3 a) {
    return a.x != 0.0f || a.y != 0.0f || a.z != 0.0f;
}
static inline int vany_nan(v3 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z); }
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static inline v3 load3(const float *p) { return V3(p[0], p[1], p[2]); }

/* safe_normalize (src/luisa/path_tracer_lighting.cpp:9-18 and
 * src/luisa/graph_surface_state.cpp:176-189). */
static v3 safe_normalize(v3 value, v3 fallback) {
    const int valid = vdot(value, value) > 1.0e-20f;
    v3 selected = vsel(fallback, value, valid);
    const int fallback_valid = vdot(selected, selected) > 1.0e-20f;
    selected = vsel(V3(0.0f, 0.0f, 1.0f), selected, fallback_valid);
    return vnormalize(selected);
}

/* normalize_or (include/psycles/luisa/spherical_geometry.h:34-40;
 * src/luisa/path_kernel_emissive_triangle.cpp:10-22). */
static v3 normalize_or(v3 value, v3 fallback) {
    const float length_squared = vdot(value, value);
    return length_squared > 1.0e-20f
               ? vdivs(value, sqrtf(fmaxf(length_squared, 1.0e-20f)))
               : fallback;
}

/* Column-major 4x4 matrix, elements[12..14] = translation (psycles::Mat4f). */
typedef struct { float m[16]; } mat4;

/* cycles_transform::point / direction (src/luisa/cycles_transform.cpp:6-27):
 * the explicit FMA tree of the Cycles affine transform. */
static v3 affine(const mat4 *t, v3 v, int translate) {
    const float *e = t->m;
    const float tx = translate ? e[12] : 0.0f;
    const float ty = translate ? e[13] : 0.0f;
    const float tz = translate ? e[14] : 0.0f;
    return V3(fmaf(v.x, e[0], fmaf(v.y, e[4], fmaf(v.z, e[8], tx))),
              fmaf(v.x, e[1], fmaf(v.y, e[5], fmaf(v.z, e[9], ty))),
              fmaf(v.x, e[2], fmaf(v.y, e[6], fmaf(v.z, e[10], tz))));
}
static inline v3 xform_point(const mat4 *t, v3 v) { return affine(t, v, 1); }
static inline v3 xform_dir(const mat4 *t, v3 v) { return affine(t, v, 0); }

/* Luisa float4x4 * float4 (columns accumulated in order, no fusion). */
static v3 mat_mul(const mat4 *t, v3 v, float w) {
    const float *e = t->m;
    v3 r;
    r.x = ((e[0] * v.x + e[4] * v.y) + e[8] * v.z) + e[12] * w;
    r.y = ((e[1] * v.x + e[5] * v.y) + e[9] * v.z) + e[13] * w;
    r.z = ((e[2] * v.x + e[6] * v.y) + e[10] * v.z) + e[14] * w;
    return r;
}

static mat4 transpose4(const mat4 *a) {
    mat4 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) r.m[c * 4 + row] = a->m[row * 4 + c];
    return r;
}

/* cycles_inverse_transform (src/luisa/path_tracer_scene_geometry.cpp:231-268). */
static mat4 cycles_inverse(const mat4 *t) {
    const float *e = t->m;
    v3 x = V3(e[0], e[1], e[2]);
    v3 y = V3(e[4], e[5], e[6]);
    v3 z = V3(e[8], e[9], e[10]);
    const v3 translation = V3(e[12], e[13], e[14]);
    float det = vdot(x, vcross(y, z));
    if (det == 0.0f) {
        x.x += 1.0e-8f;
        y.y += 1.0e-8f;
        z.z += 1.0e-8f;
        det = vdot(x, vcross(y, z));
        if (det == 0.0f) det = FLT_MAX;
    }
    const v3 ix = vdivs(vcross(y, z), det);
    const v3 iy = vdivs(vcross(z, x), det);
    const v3 iz = vdivs(vcross(x, y), det);
    mat4 r;
    const float vals[16] = {ix.x, iy.x, iz.x, 0.0f, ix.y, iy.y, iz.y, 0.0f,
                            ix.z, iy.z, iz.z, 0.0f, -vdot(ix, translation),
                            -vdot(iy, translation), -vdot(iz, translation),
                            1.0f};
    memcpy(r.m, vals, sizeof(vals));
    return r;
}

/* ======================================================================= */
/* tabulated Sobol (src/sampling/tabulated_sobol.cpp:24-199)                */
/* ======================================================================= */

static const uint32_t sobol_xors[4][32] = {
    {0u},
    {0x00000000u, 0x00000001u, 0x00000001u, 0x00000007u, 0x00000001u,
     0x00000013u, 0x00000015u, 0x0000007fu, 0x00000001u, 0x00000103u,
     0x00000105u, 0x0000070fu, 0x00000111u, 0x00001333u, 0x00001555u,
     0x00007fffu, 0x00000001u, 0x00010003u, 0x00010005u, 0x0007000fu,
     0x00010011u, 0x00130033u, 0x00150055u, 0x007f00ffu, 0x00010101u,
     0x01030303u, 0x01050505u, 0x070f0f0fu, 0x01111111u, 0x13333333u,
     0x15555555u, 0x7fffffffu},
    {0x00000000u, 0x00000001u, 0x00000003u, 0x00000001u, 0x00000005u,
     0x0000001fu, 0x0000002bu, 0x0000003du, 0x00000011u, 0x00000133u,
     0x00000377u, 0x00000199u, 0x00000445u, 0x00001ccfu, 0x00002ddbu,
     0x0000366du, 0x00000101u, 0x00010303u, 0x00030707u, 0x00010909u,
     0x00051515u, 0x001f3f3fu, 0x002b6b6bu, 0x003dbdbdu, 0x00101011u,
     0x01303033u, 0x03707077u, 0x01909099u, 0x04515145u, 0x1cf3f3cfu,
     0x2db6b6dbu, 0x36dbdb6du},
    {0x00000000u, 0x00000001u, 0x00000000u, 0x00000003u, 0x0000000du,
     0x0000000cu, 0x00000005u, 0x0000004fu, 0x00000014u, 0x000000e7u,
     0x00000329u, 0x0000039cu, 0x00000011u, 0x00001033u, 0x00000044u,
     0x000030bbu, 0x0000d1cdu, 0x0000c2ecu, 0x00005415u, 0x0004fc3fu,
     0x00015054u, 0x000e5c97u, 0x0032e5b9u, 0x0039725cu, 0x00000101u,
     0x01000303u, 0x00000404u, 0x03000b0bu, 0x0d001d1du, 0x0c002c2cu,
     0x05004545u, 0x4f00cfcfu}};

static uint32_t hash_hp_uint(uint32_t v) {
    v ^= v >> 16u;
    v *= 0x21f0aaadu;
    v ^= v >> 15u;
    v *= 0xd35a2d97u;
    v ^= v >> 15u;
    return v ^ 0xe6fe3bebu;
}

static uint32_t hash_hp_seeded_uint(uint32_t v, uint32_t seed) {
    seed ^= seed << 19u;
    return hash_hp_uint(v ^ seed);
}

static float hash_hp_float(uint32_t v) {
    return (float)hash_hp_uint(v) * (1.0f / 4294967808.0f);
}

static void sobol_generate_sequence(float *points, uint32_t size,
                                     uint32_t rng_seed) {
    uint32_t rng_index = hash_hp_seeded_uint(rng_seed, 0x44605a73u);
    for (int c = 0; c < 4; ++c) points[c] = hash_hp_float(rng_index++);
    uint32_t log_n = 0u;
    for (uint32_t n = 1u; n < size; ++log_n, n *= 2u) {
        const float strata = (float)(n * 2u);
        for (uint32_t i = 0u; i < n && n + i < size; ++i) {
            uint32_t occupied[4];
            for (int c = 0; c < 4; ++c)
                occupied[c] = (uint32_t)(points[4u * (i ^ sobol_xors[c][log_n]) + c] *
                                         strata);
            float *point = points + 4u * (n + i);
            for (int c = 0; c < 4; ++c)
                point[c] = ((float)(occupied[c] ^ 1u) + hash_hp_float(rng_index++)) /
                           strata;
        }
    }
}

uint32_t oracle_sequence_size_for_samples(uint32_t sample_count) {
    uint32_t clamped = sample_count < 1u ? 1u : sample_count;
    if (clamped > 8192u) clamped = 8192u;
    uint32_t unbounded = 1u;
    if (clamped != 1u) {
        unbounded = 1u << (32u - (uint32_t)__builtin_clz(clamped - 1u));
    }
    if (unbounded < 256u) unbounded = 256u;
    if (unbounded > 8192u) unbounded = 8192u;
    return unbounded;
}

int oracle_sobol_table(uint32_t sequence_size, float *out) {
    if (sequence_size < 256u || sequence_size > 8192u ||
        (sequence_size & (sequence_size - 1u)) != 0u)
        return -1;
    for (uint32_t pattern = 0u; pattern < 256u; ++pattern)
        sobol_generate_sequence(out + 4u * (size_t)pattern * sequence_size,
                                sequence_size, pattern);
    return 0;
}

/* ======================================================================= */
/* device sampler (include/psycles/luisa/cycles_sampler.h:36-271)           */
/* ======================================================================= */

static uint32_t reverse_bits(uint32_t v) {
    v = ((v >> 1u) & 0x55555555u) | ((v & 0x55555555u) << 1u);
    v = ((v >> 2u) & 0x33333333u) | ((v & 0x33333333u) << 2u);
    v = ((v >> 4u) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4u);
    v = ((v >> 8u) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8u);
    return (v >> 16u) | (v << 16u);
}

static uint32_t reversed_bit_owen(uint32_t v, uint32_t seed) {
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16u) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return v;
}

static uint32_t nested_uniform_scramble(uint32_t v, uint32_t seed) {
    return reverse_bits(reversed_bit_owen(reverse_bits(v), seed));
}

static uint32_t hash_wang_seeded_uint(uint32_t v, uint32_t seed) {
    v = (v ^ 61u) ^ seed;
    v += v << 3u;
    v ^= v >> 4u;
    v *= 0x27d4eb2du;
    return v;
}

static uint32_t shuffled_pattern_index(uint32_t dimension, uint32_t seed) {
    uint32_t v = dimension & 0xffu;
    const uint32_t mask = 0xffu;
    v ^= seed;
    v *= 0xe170893du;
    v ^= seed >> 16u;
    v ^= (v & mask) >> 4u;
    v ^= seed >> 8u;
    v *= 0x0929eb3fu;
    v ^= seed >> 23u;
    v ^= (v & mask) >> 1u;
    v *= 1u | (seed >> 27u);
    v *= 0x6935fa69u;
    v ^= (v & mask) >> 11u;
    v *= 0x74dcb303u;
    v ^= (v & mask) >> 2u;
    v *= 0x9e501cc3u;
    v ^= (v & mask) >> 2u;
    v *= 0xc860a3dfu;
    v &= mask;
    v ^= v >> 5u;
    return v;
}

uint32_t oracle_pixel_hash(uint32_t x, uint32_t y, uint32_t seed) {
    const uint32_t qx = 1103515245u * ((x >> 1u) ^ y);
    const uint32_t qy = 1103515245u * ((y >> 1u) ^ x);
    const uint32_t hash = 1103515245u * (qx ^ (qy >> 3u));
    return hash ^ seed;
}

uint32_t oracle_shuffled_sample_index(uint32_t sample, uint32_t dimension,
                                      uint32_t seed, uint32_t sequence_size) {
    const uint32_t pattern = shuffled_pattern_index(dimension, seed);
    const uint32_t mask = sequence_size - 1u;
    const uint32_t shuffled =
        nested_uniform_scramble(sample, hash_wang_seeded_uint(dimension, seed));
    sample = (sample & ~mask) | (shuffled & mask);
    return ((pattern * sequence_size) + sample) % (sequence_size * 256u);
}

typedef struct sampler {
    const float *table;
    uint32_t sequence_size;
} sampler;

static const float *sample_4d(const sampler *s, uint32_t sample,
                              uint32_t rng_hash, uint32_t dimension) {
    return s->table +
           4u * (size_t)oracle_shuffled_sample_index(sample, dimension, rng_hash,
                                                     s->sequence_size);
}

/* ======================================================================= */
/* pixel filter (src/sampling/pixel_filter.cpp:30-86;                       */
/*               include/psycles/luisa/pixel_filter.h:16-28)                */
/* ======================================================================= */

static float filter_eval(uint32_t filter, float value, float width) {
    if (filter == PSX_FILTER_GAUSSIAN) {
        value *= 6.0f / width;
        return expf(-2.0f * value * value);
    }
    if (filter == PSX_FILTER_BLACKMAN_HARRIS) {
        const float two_pi = 6.28318530717958647692f;
        value = two_pi * (value / width + 0.5f);
        return 0.35875f - 0.48829f * cosf(value) + 0.14128f * cosf(2.0f * value) -
               0.01168f * cosf(3.0f * value);
    }
    return 1.0f;
}

void oracle_pixel_filter_table(uint32_t filter, float width, float *out) {
    enum { size = 1024 };
    width = fmaxf(width, 1.0e-5f);
    if (filter == PSX_FILTER_GAUSSIAN) width *= 3.0f;
    else if (filter == PSX_FILTER_BLACKMAN_HARRIS) width *= 2.0f;
    const unsigned resolution = size - 1u;
    float cdf[size];
    memset(cdf, 0, sizeof(cdf));
    const float range = width * 0.5f;
    for (unsigned i = 0; i < resolution; ++i) {
        const float value = range * (float)i / (float)(resolution - 1u);
        cdf[i + 1u] = cdf[i] + fabsf(filter_eval(filter, value, width));
    }
    const float normalization =
        cdf[resolution] == 0.0f ? 0.0f : 1.0f / cdf[resolution];
    for (unsigned i = 0; i < size; ++i) cdf[i] *= normalization;
    cdf[resolution] = 1.0f;
    memset(out, 0, sizeof(float) * size);
    const unsigned half = (size - 1u) / 2u;
    for (unsigned i = 0; i <= half; ++i) {
        const float x = (float)i / (float)half;
        /* std::upper_bound: first element > x */
        unsigned index = 0;
        while (index < size && !(x < cdf[index])) ++index;
        float interpolation = 0.0f;
        if (index < resolution) {
            interpolation = (x - cdf[index]) / (cdf[index + 1u] - cdf[index]);
        } else {
            index = resolution;
        }
        const float y =
            ((float)index + interpolation) / (float)(size - 1u) * (2.0f * range);
        out[half + i] = 0.5f * (1.0f + y);
        out[half - i] = 0.5f * (1.0f - y);
    }
}

float oracle_pixel_filter_sample(const float *table, float random) {
    const unsigned last = 1023u;
    const float coordinate = fclamp(random, 0.0f, 1.0f) * (float)last;
    unsigned index = (unsigned)coordinate;
    if (index > last) index = last;
    const unsigned next = index + 1u < last ? index + 1u : last;
    const float t = coordinate - (float)index;
    return table[index] + t * (table[next] - table[index]);
}

/* ======================================================================= */
/* sample mappings (include/psycles/luisa/cycles_sample_mapping.h:49-121)  */
/* ======================================================================= */

static const float PI_F = 3.14159265358979323846f;
static const float INV_PI_F = 0.31830988618379067154f;

static v2 sample_uniform_disk(v2 random) {
    const float a = 2.0f * random.x - 1.0f;
    const float b = 2.0f * random.y - 1.0f;
    const int x_major = a * a > b * b;
    const float safe_a = a != 0.0f ? a : 1.0f;
    const float safe_b = b != 0.0f ? b : 1.0f;
    const float radius = x_major ? a : b;
    const float angle = x_major ? 0.25f * PI_F * (b / safe_a)
                                : 0.5f * PI_F - 0.25f * PI_F * (a / safe_b);
    v2 mapped = {cosf(angle) * radius, sinf(angle) * radius};
    if (a == 0.0f && b == 0.0f) mapped.x = mapped.y = 0.0f;
    return mapped;
}

static void make_orthonormals(v3 n, v3 *tangent, v3 *bitangent) {
    const v3 general = V3(n.z - n.y, n.x - n.z, n.y - n.x);
    const v3 equal = V3(n.z - n.y, n.x + n.z, -n.y - n.x);
    const int use_general = (n.x != n.y) || (n.x != n.z);
    *tangent = vnormalize(use_general ? general : equal);
    *bitangent = vcross(n, *tangent);
}

static v3 sample_cosine_hemisphere(v3 normal, v2 random, float *pdf) {
    const v2 disk = sample_uniform_disk(random);
    const float cosine = sqrtf(fmaxf(1.0f - (disk.x * disk.x + disk.y * disk.y), 0.0f));
    v3 t, b;
    make_orthonormals(normal, &t, &b);
    if (pdf) *pdf = cosine * INV_PI_F;
    return vadd(vadd(vscale(t, disk.x), vscale(b, disk.y)), vscale(normal, cosine));
}

void oracle_sample_uniform_disk(const float random[2], float out[2]) {
    v2 r = {random[0], random[1]};
    v2 d = sample_uniform_disk(r);
    out[0] = d.x;
    out[1] = d.y;
}

void oracle_sample_cosine_hemisphere(const float normal[3], const float random[2],
                                     float out[4]) {
    v2 r = {random[0], random[1]};
    float pdf;
    v3 d = sample_cosine_hemisphere(load3(normal), r, &pdf);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
    out[3] = pdf;
}

/* ======================================================================= */
/* camera helpers (include/psycles/luisa/camera_sampling.h:81-107)         */
/* ======================================================================= */

void oracle_camera_clip_range(float near_clip, float far_clip, float cosine,
                              float out[2]) {
    const float safe = fmaxf(fabsf(cosine), 1.0e-20f);
    out[0] = near_clip / safe;
    out[1] = (far_clip - near_clip) / safe;
}

void oracle_affine_point(const float m[16], const float p[3], float out[3]) {
    mat4 t;
    memcpy(t.m, m, sizeof(t.m));
    v3 r = xform_point(&t, load3(p));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

/* ======================================================================= */
/* Cycles path state (include/psycles/luisa/cycles_path_state.h)           */
/* ======================================================================= */

enum {
    FLAG_REFLECT = 1u << 0, FLAG_SINGULAR = 1u << 1, FLAG_TRANSPARENT = 1u << 2,
    FLAG_IMPORTANCE_BAKE = 1u << 3, FLAG_DIFFUSE_ANCESTOR = 1u << 4,
    FLAG_EMISSION = 1u << 5, FLAG_MIS_HAD_TRANSMISSION = 1u << 6,
    FLAG_MIS_SKIP = 1u << 7, FLAG_SINGLE_PASS_DONE = 1u << 8,
    FLAG_TRANSPARENT_BACKGROUND = 1u << 9,
    FLAG_TERMINATE_ON_NEXT_SURFACE = 1u << 10,
    FLAG_TERMINATE_IN_NEXT_VOLUME = 1u << 11,
    FLAG_TERMINATE_AFTER_TRANSPARENT = 1u << 12,
    FLAG_SURFACE_PASS = 1u << 18, FLAG_VOLUME_PASS = 1u << 19,
    FLAG_ANY_PASS = (1u << 18) | (1u << 19)
};
enum {
    CVIS_CAMERA = 1u << 0, CVIS_TRANSMIT = 1u << 1, CVIS_DIFFUSE = 1u << 2,
    CVIS_GLOSSY = 1u << 3, CVIS_VOLUME_SCATTER = 1u << 4
};
enum {
    LABEL_TRANSMIT = 1u, LABEL_REFLECT = 2u, LABEL_DIFFUSE = 4u,
    LABEL_GLOSSY = 8u, LABEL_SINGULAR = 16u, LABEL_TRANSPARENT = 32u,
    LABEL_TRANSMIT_TRANSPARENT = 128u, LABEL_SUBSURFACE = 256u,
    LABEL_RAY_PORTAL = 512u
};
enum {
    RT_BACKFACING = 1u << 0, RT_EMISSION = 1u << 2, RT_BSDF = 1u << 3,
    RT_BSDF_HAS_EVAL = 1u << 4, RT_TRANSPARENT = 1u << 10,
    RT_BSDF_HAS_TRANSMISSION = 1u << 11, RT_RAY_PORTAL = 1u << 12
};
enum {
    EV_DIFFUSE = 1u << 0, EV_GLOSSY = 1u << 1, EV_SINGULAR = 1u << 2,
    EV_REFLECTION = 1u << 3, EV_TRANSMISSION = 1u << 4,
    EV_TRANSPARENT = 1u << 5, EV_SUBSURFACE = 1u << 6
};

void oracle_initial_state(oracle_path_state *s) {
    memset(s, 0, sizeof(*s));
    s->flag = FLAG_MIS_SKIP | FLAG_TRANSPARENT_BACKGROUND;
    s->visibility = CVIS_CAMERA;
    s->rng_offset = 16u;
}

/* next_surface (cycles_path_state.h:205-362) */
void oracle_next_surface(const oracle_path_state *before, uint32_t label,
                         uint32_t runtime_flag, const oracle_path_limits *lim,
                         oracle_path_state *after) {
    const int transparent = (label & (LABEL_TRANSPARENT | LABEL_RAY_PORTAL)) != 0u;
    if (transparent) {
        oracle_path_state t = *before;
        t.transparent_bounce += 1u;
        t.rng_offset += 16u;
        t.flag |= FLAG_TRANSPARENT;
        if (t.transparent_bounce >= lim->maximum_transparent)
            t.flag |= FLAG_TERMINATE_ON_NEXT_SURFACE;
        if ((runtime_flag & RT_RAY_PORTAL) != 0u) t.flag |= FLAG_MIS_SKIP;
        *after = t;
        return;
    }
    oracle_path_state s = *before;
    s.bounce += 1u;
    s.rng_offset += 16u;
    if (s.bounce >= lim->maximum) s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    s.visibility = 0u;
    s.flag &= ~(FLAG_REFLECT | FLAG_SINGULAR | FLAG_TRANSPARENT |
                FLAG_IMPORTANCE_BAKE | FLAG_MIS_SKIP | FLAG_MIS_HAD_TRANSMISSION);
    const int reflection = (label & LABEL_REFLECT) != 0u;
    const int diffuse = (label & LABEL_DIFFUSE) != 0u;
    const int glossy = (label & LABEL_GLOSSY) != 0u;
    const int transmit_transparent = (label & LABEL_TRANSMIT_TRANSPARENT) != 0u;
    if (reflection) s.flag |= FLAG_REFLECT;
    if (reflection || !transmit_transparent) s.flag &= ~FLAG_TRANSPARENT_BACKGROUND;
    const int diffuse_reflection = reflection && diffuse;
    const int glossy_reflection = reflection && !diffuse;
    const int transmission = !reflection;
    if (diffuse_reflection) s.diffuse_bounce += 1u;
    if (glossy_reflection) s.glossy_bounce += 1u;
    if (transmission) s.transmission_bounce += 1u;
    if (diffuse_reflection && s.diffuse_bounce >= lim->maximum_diffuse)
        s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    if (glossy_reflection && s.glossy_bounce >= lim->maximum_glossy)
        s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    if (transmission && s.transmission_bounce >= lim->maximum_transmission)
        s.flag |= FLAG_TERMINATE_AFTER_TRANSPARENT;
    if (transmission) s.visibility |= CVIS_TRANSMIT;
    s.visibility |= diffuse ? CVIS_DIFFUSE : CVIS_GLOSSY;
    if (diffuse) s.flag |= FLAG_DIFFUSE_ANCESTOR;
    if (!diffuse && !glossy) s.flag |= FLAG_SINGULAR | FLAG_MIS_SKIP;
    if ((runtime_flag & RT_BSDF_HAS_TRANSMISSION) != 0u)
        s.flag |= FLAG_MIS_HAD_TRANSMISSION;
    if ((s.flag & FLAG_ANY_PASS) == 0u && (s.flag & FLAG_TRANSPARENT_BACKGROUND) == 0u)
        s.flag |= FLAG_SURFACE_PASS;
    *after = s;
}

/* contract_visibility (cycles_path_state.h:448-489) */
uint32_t oracle_contract_visibility(uint32_t v) {
    const uint32_t t = (v & CVIS_TRANSMIT) ? (v & ~(CVIS_DIFFUSE | CVIS_GLOSSY)) : v;
    uint32_t r = 0u;
    if (t & CVIS_CAMERA) r |= PSX_VISIBILITY_CAMERA;
    if (t & CVIS_TRANSMIT) r |= PSX_VISIBILITY_TRANSMISSION;
    if (t & CVIS_DIFFUSE) r |= PSX_VISIBILITY_DIFFUSE;
    if (t & CVIS_GLOSSY) r |= PSX_VISIBILITY_GLOSSY;
    if (t & CVIS_VOLUME_SCATTER) r |= PSX_VISIBILITY_VOLUME_SCATTER;
    return r;
}

/* continuation_probability (cycles_path_state.h:416-446) */
float oracle_continuation_probability(uint32_t flag, uint32_t bounce,
                                      uint32_t transparent_bounce,
                                      uint32_t minimum, uint32_t transparent_minimum,
                                      const float th[3]) {
    const int transparent = (flag & FLAG_TRANSPARENT) != 0u;
    const int before_minimum = transparent ? transparent_bounce <= transparent_minimum
                                           : bounce <= minimum;
    const float roulette = fminf(
        sqrtf(fmaxf(fabsf(th[0]), fmaxf(fabsf(th[1]), fabsf(th[2])))), 1.0f);
    return before_minimum ? 1.0f : roulette;
}

static uint32_t label_from_events(uint32_t e) {
    uint32_t l = 0u;
    if (e & EV_TRANSMISSION) l |= LABEL_TRANSMIT;
    if (e & EV_REFLECTION) l |= LABEL_REFLECT;
    if (e & EV_DIFFUSE) l |= LABEL_DIFFUSE;
    if (e & EV_GLOSSY) l |= LABEL_GLOSSY;
    if (e & EV_SINGULAR) l |= LABEL_SINGULAR;
    if (e & EV_TRANSPARENT) l |= LABEL_TRANSPARENT;
    if (e & EV_SUBSURFACE) l |= LABEL_SUBSURFACE;
    return l;
}

/* ======================================================================= */
/* light distribution (src/sampling/light_distribution.cpp:25-138)         */
/* ======================================================================= */

typedef struct dist_entry {
    float cumulative;
    float selection_pdf;
    uint32_t kind;  /* 0 triangle, 1 analytic, 2 environment, 3 sentinel */
    uint32_t index;
} dist_entry;

static float valid_area(float a) { return (isfinite(a) && a > 0.0f) ? a : 0.0f; }

static uint32_t build_distribution(const float *areas, uint32_t tri_count,
                                   uint32_t analytic_count, int include_env,
                                   dist_entry *e, float *tri_pdf,
                                   float *light_pdf) {
    const uint32_t env_count = include_env ? 1u : 0u;
    const uint32_t light_count = analytic_count + env_count;
    const uint32_t emitter_count = tri_count + light_count;
    uint32_t n = 0;
    float total = 0.0f;
    for (uint32_t i = 0; i < tri_count; ++i) {
        e[n].cumulative = total; e[n].selection_pdf = 0.0f; e[n].kind = 0u; e[n].index = i;
        ++n;
        total += valid_area(areas[i]);
    }
    const float triangle_area = total;
    const float light_area =
        light_count > 0u ? (total > 0.0f ? total / (float)light_count : 1.0f) : 0.0f;
    for (uint32_t i = 0; i < analytic_count; ++i) {
        e[n].cumulative = total; e[n].selection_pdf = 0.0f; e[n].kind = 1u; e[n].index = i;
        ++n;
        total += light_area;
    }
    if (include_env) {
        e[n].cumulative = total; e[n].selection_pdf = 0.0f; e[n].kind = 2u; e[n].index = 0u;
        ++n;
        total += light_area;
    }
    e[n].cumulative = total; e[n].selection_pdf = 0.0f; e[n].kind = 3u; e[n].index = 0u;
    float tp = 0.0f, lp = 0.0f;
    if (triangle_area > 0.0f) {
        tp = 1.0f / triangle_area;
        if (light_count > 0u) tp *= 0.5f;
    }
    if (light_count > 0u) {
        lp = 1.0f / (float)light_count;
        if (triangle_area > 0.0f) lp *= 0.5f;
    }
    for (uint32_t i = 0; i < tri_count; ++i) e[i].selection_pdf = valid_area(areas[i]) * tp;
    for (uint32_t i = tri_count; i < emitter_count; ++i) e[i].selection_pdf = lp;
    if (total > 0.0f) {
        const float inv = 1.0f / total;
        for (uint32_t i = 0; i < emitter_count; ++i) e[i].cumulative *= inv;
        e[emitter_count].cumulative = 1.0f;
    }
    *tri_pdf = tp;
    *light_pdf = lp;
    return emitter_count;
}

uint32_t oracle_light_distribution(const float *areas, uint32_t tri_count,
                                   uint32_t analytic_count, int include_env,
                                   uint32_t *entries, float *tri_pdf,
                                   float *light_pdf) {
    dist_entry *e = (dist_entry *)malloc(sizeof(dist_entry) *
                                         (tri_count + analytic_count + 2u));
    const uint32_t n = build_distribution(areas, tri_count, analytic_count,
                                          include_env, e, tri_pdf, light_pdf);
    for (uint32_t i = 0; i <= n; ++i) {
        entries[4 * i + 0] = fbits(e[i].cumulative);
        entries[4 * i + 1] = fbits(e[i].selection_pdf);
        entries[4 * i + 2] = e[i].kind;
        entries[4 * i + 3] = e[i].index;
    }
    free(e);
    return n;
}

/* Upper-bound search (src/luisa/path_tracer_light_distribution.cpp:12-52). */
static const dist_entry *distribution_sample(const dist_entry *e, uint32_t count,
                                             float u) {
    uint32_t first = 0u, length = count + 1u;
    while (length > 0u) {
        const uint32_t half = length >> 1u;
        const uint32_t middle = first + half;
        if (u < e[middle].cumulative) {
            length = half;
        } else {
            first = middle + 1u;
            length = length - half - 1u;
        }
    }
    uint32_t selected = first > 0u ? first - 1u : 0u;
    if (selected > count - 1u) selected = count - 1u;
    return &e[selected];
}

/* ======================================================================= */
/* Pluecker triangle test                                                   */
/* (src/luisa/cycles_triangle_intersection_component.cpp:14-108)           */
/* ======================================================================= */

static v3 pluecker_cross(v3 a, v3 b) {
    return V3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
              fmaf(a.x, b.y, -(a.y * b.x)));
}
static float pluecker_dot(v3 a, v3 b) {
    return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
}
static float clamp_bvh_direction(float v) {
    const float m = 8.271806e-25f;
    return fabsf(v) > m ? v : copysignf(m, v);
}

typedef struct tri_hit {
    int valid;
    float t, u, v;
} tri_hit;

/* origin/direction are already in the intersection space of the triangle. */
static tri_hit pluecker(v3 origin, v3 raw_direction, float tmin, float tmax,
                        v3 p0, v3 p1, v3 p2) {
    const v3 d = V3(clamp_bvh_direction(raw_direction.x),
                    clamp_bvh_direction(raw_direction.y),
                    clamp_bvh_direction(raw_direction.z));
    const v3 v0 = vsub(p0, origin), v1 = vsub(p1, origin), v2 = vsub(p2, origin);
    const v3 e0 = vsub(v2, v0), e1 = vsub(v0, v1), e2 = vsub(v1, v2);
    const float u = pluecker_dot(pluecker_cross(e0, vadd(v2, v0)), d);
    const float v = pluecker_dot(pluecker_cross(e1, vadd(v0, v1)), d);
    const float w = pluecker_dot(pluecker_cross(e2, vadd(v1, v2)), d);
    const float sum = u + v + w;
    const float tol = FLT_EPSILON * fabsf(sum);
    const float mn = fminf(u, fminf(v, w));
    const float mx = fmaxf(u, fmaxf(v, w));
    const int inside = (mn >= -tol) || (mx <= tol);
    const v3 n2 = vscale(pluecker_cross(e1, e0), 2.0f);
    const float den = pluecker_dot(n2, d);
    const float safe_den = den != 0.0f ? den : 1.0f;
    const float t = pluecker_dot(v0, n2) / safe_den;
    const int in_range = (t >= tmin) && (t <= tmax);
    const float rs = fabsf(sum) >= 1.0e-18f ? 1.0f / sum : 0.0f;
    tri_hit h;
    h.valid = inside && den != 0.0f && in_range;
    h.t = t;
    h.u = fminf(u * rs, 1.0f);
    h.v = fminf(v * rs, 1.0f);
    return h;
}

int oracle_pluecker_intersect(const float origin[3], const float direction[3],
                              float tmin, float tmax, const float p0[3],
                              const float p1[3], const float p2[3], float out[3]) {
    tri_hit h = pluecker(load3(origin), load3(direction), tmin, tmax, load3(p0),
                         load3(p1), load3(p2));
    out[0] = h.t;
    out[1] = h.u;
    out[2] = h.v;
    return h.valid;
}

/* ======================================================================= */
/* scene                                                                    */
/* ======================================================================= */

enum { KIND_DIFFUSE = 1, KIND_TRANSPARENT = 2 };

typedef struct omaterial {
    psx_closure_instr *instr;
    uint32_t count, root;
    float *params;
    uint32_t param_count;
    uint32_t emission_sampling;
    int may_emit;
} omaterial;

typedef struct bvh_node {
    float lo[3], hi[3];
    uint32_t first;  /* left child index for inner nodes, first prim for leaves */
    uint32_t count;  /* 0 for inner nodes */
} bvh_node;

typedef struct blas {
    uint32_t mesh;
    int world;              /* positions are final world support positions */
    float *ipos;            /* intersection positions, 3 * V                */
    bvh_node *nodes;
    uint32_t node_count;
    uint32_t *prims;        /* leaf order -> local primitive                */
    float root_lo[3], root_hi[3];
} blas;

typedef struct oinstance {
    uint32_t mesh, blas, visibility;
    int applied;
    mat4 o2w, w2o, n2w;
    uint32_t *override;
    uint32_t override_count;
    float terminator_offset; /* shadow_terminator_geometry_offset */
    float lo[3], hi[3];     /* world bounds */
} oinstance;

typedef struct omesh {
    const float *pos;       /* borrowed copies below */
    float *positions;
    float *normals;
    uint32_t *tris;
    uint8_t *smooth;
    uint32_t *slots;
    uint32_t slot_count;
    uint32_t *tri_slots;
    uint32_t vcount, tcount;
    uint32_t prim_offset;
    uint32_t users;
} omesh;

typedef struct emitter {
    uint32_t instance, prim;
    uint32_t material;
    uint32_t emission_sampling;
} emitter;

/* Analytic lamp (LightGpu, src/luisa/path_tracer_types.h:142-168, built by
 * src/luisa/path_tracer_scene.cpp:1723-1870): point, spot or distant. */
#define LAMP_NORMALIZE 1u
#define LAMP_ELLIPSE 2u
#define LAMP_SPHERE 4u
#define LAMP_USE_MIS 8u
#define LAMP_FORWARD 32u
#define LAMP_FULL_SPREAD 64u
#define LAMP_CAST_SHADOW 128u

typedef struct olamp {
    uint32_t type;
    v3 position, axis_x, axis_y, axis_z, axis_scale, color;
    float power, radius, spot_angle, spot_smooth, angle;
    float size_u, size_v, spread;   /* area lamps: scaled edge lengths, spread angle */
    uint32_t flags, visibility, max_bounces;
} olamp;

struct oracle_scene {
    olamp *lamps;
    uint32_t lamp_count;
    oracle_light_tree tree;     /* light tree over emissive triangles, then lamps */
    uint32_t *emitter_cprim;    /* Cycles primitive of each emissive triangle     */
    omaterial *materials;
    uint32_t material_count;
    omesh *meshes;
    uint32_t mesh_count;
    oinstance *instances;
    uint32_t instance_count;
    blas *blases;
    uint32_t blas_count;
    bvh_node *tlas;
    uint32_t tlas_count;
    uint32_t *tlas_prims;
    emitter *emitters;
    uint32_t emitter_count;
    dist_entry *distribution;
    uint32_t distribution_count;
    float triangle_area_pdf;
    float light_selection_pdf;
    psx_camera camera;
    v3 background;
    uint32_t world_visibility;
    int brute_force;
};

/* ---- materials: closure-tree flattening ------------------------------- */
/* (src/luisa/graph_surface_closures.cpp:6-353,
 *  src/luisa/graph_surface_physical_closures.cpp:200-470,
 *  src/luisa/graph_surface_state.cpp:20-58) */

typedef struct closure_rec {
    uint32_t kind;
    v3 weight;
    float sample_weight;
    v3 normal;
    float roughness;
    v3 color;
} closure_rec;

#define MAX_CLOSURES 16

typedef struct closure_set {
    closure_rec c[MAX_CLOSURES];
    uint32_t count;
    v3 emission;
} closure_set;

static float avg3(v3 v) { return (v.x + v.y + v.z) / 3.0f; }
static float sample_weight_of(v3 v) { return fabsf(avg3(v)); }

static float param(const omaterial *m, const float *block, uint32_t i) {
    (void)m;
    return block[i];
}
static v3 param3(const omaterial *m, const float *block, uint32_t i) {
    (void)m;
    return V3(block[i], block[i + 1], block[i + 2]);
}

static void flatten(const omaterial *m, uint32_t id, float mix_weight,
                    v3 shading_normal, closure_set *out) {
    const psx_closure_instr *ins = &m->instr[id];
    const float *block = m->params;
    switch (ins->op) {
    case PSX_CLOSURE_ADD:
        flatten(m, ins->a, mix_weight, shading_normal, out);
        flatten(m, ins->b, mix_weight, shading_normal, out);
        return;
    case PSX_CLOSURE_MIX: {
        const float f = fclamp(param(m, block, ins->factor), 0.0f, 1.0f);
        flatten(m, ins->a, mix_weight * (1.0f - f), shading_normal, out);
        flatten(m, ins->b, mix_weight * f, shading_normal, out);
        return;
    }
    case PSX_CLOSURE_DIFFUSE: {
        const v3 color = param3(m, block, ins->color);
        /* bsdf_allocated_weight */
        v3 w = vmax(vscale(color, mix_weight), V3(0, 0, 0));
        if (!(avg3(w) >= 1.0e-5f)) w = V3(0, 0, 0);
        /* physical diffuse record */
        const float allocation = sample_weight_of(vmax(w, V3(0, 0, 0)));
        const int allocated = allocation >= 1.0e-5f;
        if (allocated && out->count < MAX_CLOSURES) {
            closure_rec *c = &out->c[out->count++];
            c->kind = KIND_DIFFUSE;
            c->weight = vmax(w, V3(0, 0, 0));
            c->sample_weight = allocation;
            c->normal = safe_normalize(shading_normal, shading_normal);
            c->roughness = param(m, block, ins->roughness);
            c->color = color;
        }
        return;
    }
    case PSX_CLOSURE_EMISSION: {
        const v3 color = param3(m, block, ins->color);
        const float strength = param(m, block, ins->strength);
        out->emission = vadd(out->emission, vscale(vscale(color, strength), mix_weight));
        return;
    }
    case PSX_CLOSURE_TRANSPARENT: {
        const v3 color = param3(m, block, ins->color);
        const v3 w = vscale(color, mix_weight);
        const float sw = sample_weight_of(w);
        if (sw >= 1.0e-5f && out->count < MAX_CLOSURES) {
            closure_rec *c = &out->c[out->count++];
            c->kind = KIND_TRANSPARENT;
            c->weight = w;
            c->sample_weight = sw;
            c->normal = shading_normal;
            c->roughness = 0.0f;
            c->color = V3(1, 1, 1);
        }
        return;
    }
    default:
        return;
    }
}

static void material_closures(const omaterial *m, v3 shading_normal, closure_set *out) {
    out->count = 0;
    out->emission = V3(0, 0, 0);
    if (m->count == 0) return;
    flatten(m, m->root, 1.0f, shading_normal, out);
}

/* constant_emission (src/luisa/graph_surface_emission.cpp:33-80): one running
 * sum, added to in tree-visit order. */
static void constant_emission_rec(const omaterial *m, uint32_t id, float w, v3 *sum) {
    const psx_closure_instr *ins = &m->instr[id];
    switch (ins->op) {
    case PSX_CLOSURE_EMISSION:
        *sum = vadd(*sum, vscale(vscale(param3(m, m->params, ins->color),
                                        param(m, m->params, ins->strength)), w));
        return;
    case PSX_CLOSURE_ADD:
        constant_emission_rec(m, ins->a, w, sum);
        constant_emission_rec(m, ins->b, w, sum);
        return;
    case PSX_CLOSURE_MIX: {
        const float f = fclamp(param(m, m->params, ins->factor), 0.0f, 1.0f);
        constant_emission_rec(m, ins->a, w * (1.0f - f), sum);
        constant_emission_rec(m, ins->b, w * f, sum);
        return;
    }
    default:
        return;
    }
}

static v3 constant_emission(const omaterial *m) {
    v3 sum = V3(0, 0, 0);
    if (m->count) constant_emission_rec(m, m->root, 1.0f, &sum);
    return sum;
}

/* estimate_surface_emission (src/compiler/surface_program.cpp:244-330) */
static v3 emission_estimate(const omaterial *m, uint32_t id) {
    const psx_closure_instr *ins = &m->instr[id];
    switch (ins->op) {
    case PSX_CLOSURE_EMISSION:
        return vscale(param3(m, m->params, ins->color), param(m, m->params, ins->strength));
    case PSX_CLOSURE_ADD:
        return vadd(emission_estimate(m, ins->a), emission_estimate(m, ins->b));
    case PSX_CLOSURE_MIX: {
        const float f = param(m, m->params, ins->factor);
        return vadd(vscale(emission_estimate(m, ins->a), 1.0f - f),
                    vscale(emission_estimate(m, ins->b), f));
    }
    default:
        return V3(0, 0, 0);
    }
}

static void material_refresh(omaterial *m) {
    v3 e = m->count ? emission_estimate(m, m->root) : V3(0, 0, 0);
    m->may_emit = (e.x != 0.0f || e.y != 0.0f || e.z != 0.0f) &&
                  m->emission_sampling != PSX_EMISSION_SAMPLING_NONE;
}

/* ---- BVH (binned SAH, 16 bins, leaf <= 4) ------------------------------ */

typedef struct build_item {
    float lo[3], hi[3], c[3];
    uint32_t id;
} build_item;

typedef struct node_vec {
    bvh_node *data;
    uint32_t size, cap;
} node_vec;

static uint32_t node_push(node_vec *v) {
    if (v->size == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->data = (bvh_node *)realloc(v->data, sizeof(bvh_node) * v->cap);
    }
    memset(&v->data[v->size], 0, sizeof(bvh_node));
    return v->size++;
}

static float box_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

static void pad_box(float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        const float ext = hi[a] - lo[a];
        const float s = fmaxf(fmaxf(fabsf(lo[a]), fabsf(hi[a])), ext);
        const float pad = s * 1.0e-5f + 1.0e-30f;
        lo[a] -= pad;
        hi[a] += pad;
    }
}

static void build_rec(node_vec *nodes, uint32_t node, build_item *items,
                      uint32_t begin, uint32_t end, uint32_t *order) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = begin; i < end; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], items[i].lo[a]);
            hi[a] = fmaxf(hi[a], items[i].hi[a]);
            clo[a] = fminf(clo[a], items[i].c[a]);
            chi[a] = fmaxf(chi[a], items[i].c[a]);
        }
    pad_box(lo, hi);
    memcpy(nodes->data[node].lo, lo, sizeof(lo));
    memcpy(nodes->data[node].hi, hi, sizeof(hi));
    const uint32_t n = end - begin;
    if (n <= 4u) {
        nodes->data[node].first = begin;
        nodes->data[node].count = n;
        for (uint32_t i = begin; i < end; ++i) order[i] = items[i].id;
        return;
    }
    enum { BINS = 16 };
    float best_cost = FLT_MAX;
    int best_axis = -1;
    int best_split = 0;
    for (int a = 0; a < 3; ++a) {
        const float extent = chi[a] - clo[a];
        if (!(extent > 0.0f)) continue;
        uint32_t cnt[BINS] = {0};
        float blo[BINS][3], bhi[BINS][3];
        for (int b = 0; b < BINS; ++b)
            for (int k = 0; k < 3; ++k) { blo[b][k] = FLT_MAX; bhi[b][k] = -FLT_MAX; }
        const float scale = (float)BINS / extent;
        for (uint32_t i = begin; i < end; ++i) {
            int b = (int)((items[i].c[a] - clo[a]) * scale);
            if (b >= BINS) b = BINS - 1;
            if (b < 0) b = 0;
            cnt[b]++;
            for (int k = 0; k < 3; ++k) {
                blo[b][k] = fminf(blo[b][k], items[i].lo[k]);
                bhi[b][k] = fmaxf(bhi[b][k], items[i].hi[k]);
            }
        }
        float left_area[BINS], right_area[BINS];
        uint32_t left_count[BINS], right_count[BINS];
        float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        uint32_t c = 0;
        for (int b = 0; b < BINS; ++b) {
            c += cnt[b];
            for (int k = 0; k < 3; ++k) { l[k] = fminf(l[k], blo[b][k]); h[k] = fmaxf(h[k], bhi[b][k]); }
            left_area[b] = box_area(l, h);
            left_count[b] = c;
        }
        for (int k = 0; k < 3; ++k) { l[k] = FLT_MAX; h[k] = -FLT_MAX; }
        c = 0;
        for (int b = BINS - 1; b >= 0; --b) {
            c += cnt[b];
            for (int k = 0; k < 3; ++k) { l[k] = fminf(l[k], blo[b][k]); h[k] = fmaxf(h[k], bhi[b][k]); }
            right_area[b] = box_area(l, h);
            right_count[b] = c;
        }
        for (int b = 0; b < BINS - 1; ++b) {
            if (left_count[b] == 0 || right_count[b + 1] == 0) continue;
            const float cost = left_area[b] * (float)left_count[b] +
                               right_area[b + 1] * (float)right_count[b + 1];
            if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = b; }
        }
    }
    uint32_t mid;
    if (best_axis < 0) {
        mid = begin + n / 2u;
    } else {
        const float extent = chi[best_axis] - clo[best_axis];
        const float scale = (float)BINS / extent;
        uint32_t i = begin, j = end;
        while (i < j) {
            int b = (int)((items[i].c[best_axis] - clo[best_axis]) * scale);
            if (b >= BINS) b = BINS - 1;
            if (b < 0) b = 0;
            if (b <= best_split) {
                ++i;
            } else {
                --j;
                build_item t = items[i]; items[i] = items[j]; items[j] = t;
            }
        }
        mid = i;
        if (mid == begin || mid == end) mid = begin + n / 2u;
    }
    const uint32_t left = node_push(nodes);
    const uint32_t right = node_push(nodes);
    (void)right;
    nodes->data[node].first = left;
    nodes->data[node].count = 0;
    build_rec(nodes, left, items, begin, mid, order);
    build_rec(nodes, left + 1u, items, mid, end, order);
}

static void build_bvh(build_item *items, uint32_t n, bvh_node **out_nodes,
                      uint32_t *out_count, uint32_t *order) {
    node_vec nodes = {0};
    const uint32_t root = node_push(&nodes);
    if (n == 0) {
        nodes.data[root].lo[0] = nodes.data[root].lo[1] = nodes.data[root].lo[2] = 1.0f;
        nodes.data[root].hi[0] = nodes.data[root].hi[1] = nodes.data[root].hi[2] = -1.0f;
        nodes.data[root].count = 0;
        nodes.data[root].first = 0;
    } else {
        build_rec(&nodes, root, items, 0, n, order);
    }
    *out_nodes = nodes.data;
    *out_count = nodes.size;
}

/* ---- scene creation -------------------------------------------------- */

static void tri_positions(const float *pos, const uint32_t *tris, uint32_t prim,
                          v3 *p0, v3 *p1, v3 *p2) {
    *p0 = load3(pos + 3u * tris[3u * prim + 0]);
    *p1 = load3(pos + 3u * tris[3u * prim + 1]);
    *p2 = load3(pos + 3u * tris[3u * prim + 2]);
}

static void build_blas(blas *b, const omesh *m) {
    build_item *items = (build_item *)malloc(sizeof(build_item) * (m->tcount ? m->tcount : 1));
    for (uint32_t t = 0; t < m->tcount; ++t) {
        v3 p0, p1, p2;
        tri_positions(b->ipos, m->tris, t, &p0, &p1, &p2);
        items[t].lo[0] = fminf(p0.x, fminf(p1.x, p2.x));
        items[t].lo[1] = fminf(p0.y, fminf(p1.y, p2.y));
        items[t].lo[2] = fminf(p0.z, fminf(p1.z, p2.z));
        items[t].hi[0] = fmaxf(p0.x, fmaxf(p1.x, p2.x));
        items[t].hi[1] = fmaxf(p0.y, fmaxf(p1.y, p2.y));
        items[t].hi[2] = fmaxf(p0.z, fmaxf(p1.z, p2.z));
        for (int a = 0; a < 3; ++a) items[t].c[a] = 0.5f * (items[t].lo[a] + items[t].hi[a]);
        items[t].id = t;
    }
    b->prims = (uint32_t *)malloc(sizeof(uint32_t) * (m->tcount ? m->tcount : 1));
    build_bvh(items, m->tcount, &b->nodes, &b->node_count, b->prims);
    memcpy(b->root_lo, b->nodes[0].lo, sizeof(b->root_lo));
    memcpy(b->root_hi, b->nodes[0].hi, sizeof(b->root_hi));
    free(items);
}

static uint32_t material_of(const oracle_scene *s, const oinstance *inst, uint32_t prim) {
    const omesh *m = &s->meshes[inst->mesh];
    const uint32_t slot = m->tri_slots ? m->tri_slots[prim] : 0u;
    if (slot < inst->override_count) return inst->override[slot];
    if (m->slot_count == 0) return UINT32_MAX;
    return m->slots[slot < m->slot_count ? slot : m->slot_count - 1u];
}

static void *dupmem(const void *src, size_t bytes) {
    if (!src || bytes == 0) return NULL;
    void *p = malloc(bytes);
    memcpy(p, src, bytes);
    return p;
}

/* ======================================================================= */
/* light tree: emitter measures and device-side selection                  */
/* (src/luisa/path_tracer_light_tree_scene.cpp:130-262,                    */
/*  src/luisa/path_tracer_light_tree.cpp:14-826)                           */
/* ======================================================================= */
#define LT_PI 3.14159265358979323846f
#define LT_HALF_PI (0.5f * LT_PI)
#define LT_SAFE 1.0e-20f
#define LT_INVALID 0xffffffffu

static v3 lt_normalize_or(v3 v, v3 fallback) {
    const float m = sqrtf(fmaxf(vdot(v, v), 0.0f));
    return m > 1.0e-20f ? vscale(v, 1.0f / m) : fallback;
}
static float lt_average_abs(v3 v) { return (fabsf(v.x) + fabsf(v.y) + fabsf(v.z)) / 3.0f; }

/* make_triangle_light_tree_emitter (:130-178) */
static oracle_tree_emitter triangle_tree_emitter(uint32_t id, const oinstance *inst, v3 p0, v3 p1, v3 p2,
                                                 v3 estimate, uint32_t emission_sampling) {
    const v3 v0 = xform_point(&inst->o2w, p0), v1 = xform_point(&inst->o2w, p1), v2 = xform_point(&inst->o2w, p2);
    const v3 cr = vcross(vsub(v1, v0), vsub(v2, v0));
    const float area = 0.5f * sqrtf(fmaxf(vdot(cr, cr), 0.0f));
    const int front = emission_sampling == PSX_EMISSION_SAMPLING_FRONT;
    const int back = emission_sampling == PSX_EMISSION_SAMPLING_BACK;
    v3 axis = (front || back) ? lt_normalize_or(cr, V3(0, 0, 1)) : lt_normalize_or(vsub(v0, v1), V3(1, 0, 0));
    if (back) axis = vscale(axis, -1.0f);
    if (front || back) {
        const v3 c0 = V3(inst->o2w.m[0], inst->o2w.m[1], inst->o2w.m[2]);
        const v3 c1 = V3(inst->o2w.m[4], inst->o2w.m[5], inst->o2w.m[6]);
        const v3 c2 = V3(inst->o2w.m[8], inst->o2w.m[9], inst->o2w.m[10]);
        if (vdot(c0, vcross(c1, c2)) < 0.0f) axis = vscale(axis, -1.0f);
    }
    oracle_tree_emitter e;
    memset(&e, 0, sizeof(e));
    const v3 lo = V3(fminf(fminf(v0.x, v1.x), v2.x), fminf(fminf(v0.y, v1.y), v2.y), fminf(fminf(v0.z, v1.z), v2.z));
    const v3 hi = V3(fmaxf(fmaxf(v0.x, v1.x), v2.x), fmaxf(fmaxf(v0.y, v1.y), v2.y), fmaxf(fmaxf(v0.z, v1.z), v2.z));
    e.bounds_min[0] = lo.x; e.bounds_min[1] = lo.y; e.bounds_min[2] = lo.z;
    e.bounds_max[0] = hi.x; e.bounds_max[1] = hi.y; e.bounds_max[2] = hi.z;
    e.bounds_empty = 0;
    e.axis[0] = axis.x; e.axis[1] = axis.y; e.axis[2] = axis.z;
    e.theta_o = (front || back) ? 0.0f : LT_HALF_PI;
    e.theta_e = LT_HALF_PI;
    e.orientation_empty = 0;
    e.energy = area * lt_average_abs(estimate);
    const v3 c = vscale(vadd(vadd(v0, v1), v2), 1.0f / 3.0f);
    e.centroid[0] = c.x; e.centroid[1] = c.y; e.centroid[2] = c.z;
    e.emitter_id = id;
    e.distant = 0;
    return e;
}

/* make_analytic_light_tree_emitter (:180-250), lamps without shaders */
static oracle_tree_emitter lamp_tree_emitter(uint32_t id, const olamp *l) {
    v3 strength = vscale(l->color, l->power);
    if ((l->flags & LAMP_NORMALIZE) == 0u) {
        float area = 1.0f;
        if (l->type == PSX_LIGHT_POINT || l->type == PSX_LIGHT_SPOT)
            area = l->radius > 0.0f ? 4.0f * LT_PI * l->radius * l->radius : 4.0f;
        else if (l->type == PSX_LIGHT_AREA) {
            area = fabsf(l->size_u * l->size_v);
            if ((l->flags & LAMP_ELLIPSE) != 0u) area *= 0.25f * LT_PI;
        } else if (l->type == PSX_LIGHT_DISTANT && l->angle > 0.0f) {
            const float s = sinf(0.5f * l->angle);
            area = LT_PI * s * s;
        }
        strength = vscale(strength, area);
    }
    oracle_tree_emitter e;
    memset(&e, 0, sizeof(e));
    e.bounds_empty = 1;
    const v3 axis = vscale(l->axis_z, -1.0f);
    e.axis[0] = axis.x; e.axis[1] = axis.y; e.axis[2] = axis.z;
    e.orientation_empty = 0;
    if (l->type == PSX_LIGHT_AREA) {
        e.theta_e = 0.5f * l->spread;
        const v3 hu = vscale(l->axis_x, 0.5f * l->size_u), hv = vscale(l->axis_y, 0.5f * l->size_v);
        const v3 c[4] = {vadd(l->position, vadd(hu, hv)), vadd(l->position, vsub(hu, hv)),
                         vadd(l->position, vsub(hv, hu)), vsub(l->position, vadd(hu, hv))};
        e.bounds_min[0] = e.bounds_max[0] = c[0].x;
        e.bounds_min[1] = e.bounds_max[1] = c[0].y;
        e.bounds_min[2] = e.bounds_max[2] = c[0].z;
        for (int i = 1; i < 4; ++i) {
            e.bounds_min[0] = fminf(e.bounds_min[0], c[i].x); e.bounds_max[0] = fmaxf(e.bounds_max[0], c[i].x);
            e.bounds_min[1] = fminf(e.bounds_min[1], c[i].y); e.bounds_max[1] = fmaxf(e.bounds_max[1], c[i].y);
            e.bounds_min[2] = fminf(e.bounds_min[2], c[i].z); e.bounds_max[2] = fmaxf(e.bounds_max[2], c[i].z);
        }
        e.bounds_empty = 0;
        strength = vscale(strength, 1.0f / LT_PI);
    } else if (l->type == PSX_LIGHT_POINT) {
        e.theta_o = LT_PI;
        e.theta_e = LT_HALF_PI;
    } else if (l->type == PSX_LIGHT_SPOT) {
        float theta_e = fminf(0.5f * l->spot_angle, LT_HALF_PI);
        if (fabsf(LT_HALF_PI - theta_e) >= 1.0e-6f)
            theta_e = atanf(tanf(theta_e) * fmaxf(l->axis_scale.x, l->axis_scale.y) / fmaxf(l->axis_scale.z, 1.0e-20f));
        e.theta_e = theta_e;
    } else {
        e.theta_e = 0.5f * l->angle;
        e.distant = 1;
    }
    if (l->type == PSX_LIGHT_POINT || l->type == PSX_LIGHT_SPOT) {
        const v3 r = V3(l->radius, l->radius, l->radius);
        const v3 a = vsub(l->position, r), b = vadd(l->position, r);
        e.bounds_min[0] = fminf(a.x, b.x); e.bounds_min[1] = fminf(a.y, b.y); e.bounds_min[2] = fminf(a.z, b.z);
        e.bounds_max[0] = fmaxf(a.x, b.x); e.bounds_max[1] = fmaxf(a.y, b.y); e.bounds_max[2] = fmaxf(a.z, b.z);
        e.bounds_empty = 0;
        strength = vscale(strength, 0.25f / LT_PI);
    }
    e.energy = lt_average_abs(strength);
    e.centroid[0] = l->position.x; e.centroid[1] = l->position.y; e.centroid[2] = l->position.z;
    e.emitter_id = id;
    return e;
}

typedef struct lt_measure {
    v3 bmin, bmax, axis;
    float theta_o, theta_e, energy;
    int has_bounds, has_orientation, distant;
} lt_measure;

static lt_measure lt_unpack(const float *m, const uint32_t *flags, int distant) {
    lt_measure r;
    r.bmin = V3(m[0], m[1], m[2]);
    r.bmax = V3(m[3], m[4], m[5]);
    r.axis = V3(m[6], m[7], m[8]);
    r.theta_o = m[9];
    r.theta_e = m[10];
    r.energy = m[11];
    r.has_bounds = flags[0] == 0u;
    r.has_orientation = flags[1] == 0u;
    r.distant = distant;
    return r;
}

static float lt_sin_from_cos(float c) { return sqrtf(fmaxf(1.0f - c * c, 0.0f)); }

/* safe_normalize of path_tracer_light_tree.cpp:19-27 */
static v3 lt_safe_normalize(v3 v, v3 fallback) {
    const float l2 = vdot(v, v);
    return l2 > LT_SAFE ? vscale(v, 1.0f / sqrtf(fmaxf(l2, LT_SAFE))) : fallback;
}

/* cos_bound_subtended_angle (:40-53) */
static float lt_cos_bound(v3 bmax, v3 centroid, v3 point) {
    const v3 pc = vsub(point, centroid), bc = vsub(bmax, centroid);
    const float d2 = vdot(pc, pc), r2 = vdot(bc, bc);
    return d2 <= r2 ? -1.0f : sqrtf(fmaxf(1.0f - r2 / fmaxf(d2, LT_SAFE), 0.0f));
}

typedef struct lt_imp { float max, min; } lt_imp;

/* bounded_importance (:83-180), surface variant */
static lt_imp lt_bounded(v3 n, int has_transmission, v3 p2c, float cos_theta_u, v3 cone_axis, float theta_o,
                         float theta_e, float maximum_distance, float minimum_distance, float energy) {
    cos_theta_u = fclamp(cos_theta_u, -1.0f, 1.0f);
    const float sin_theta_u = lt_sin_from_cos(cos_theta_u);
    int visible = energy > 0.0f;
    const float ic = has_transmission ? fabsf(vdot(p2c, n)) : vdot(p2c, n);
    const float is = lt_sin_from_cos(fclamp(ic, -1.0f, 1.0f));
    const float minimum_incidence = ic >= cos_theta_u ? 1.0f : ic * cos_theta_u + is * sin_theta_u;
    const float maximum_incidence = fmaxf(ic * cos_theta_u - is * sin_theta_u, 0.0f);
    visible = visible && (has_transmission || minimum_incidence >= 0.0f);
    const float tc = fclamp(vdot(cone_axis, vneg(p2c)), -1.0f, 1.0f);
    const float ts = lt_sin_from_cos(tc);
    const float tmu_c = tc * cos_theta_u + ts * sin_theta_u;
    const float to_c = cosf(theta_o), to_s = sinf(theta_o);
    const int fully = (tc >= cos_theta_u) || (tmu_c >= to_c);
    const int partially = !fully && ((theta_o + theta_e > LT_PI) || (tmu_c > cosf(theta_o + theta_e)));
    const float tmu_s = lt_sin_from_cos(fclamp(tmu_c, -1.0f, 1.0f));
    const float minimum_outgoing = fully ? 1.0f : (partially ? tmu_c * to_c + tmu_s * to_s : 0.0f);
    visible = visible && (fully || partially);
    const float mind = fmaxf(minimum_distance, LT_SAFE);
    const float maxd = fmaxf(maximum_distance, LT_SAFE);
    const float geometric = 1.0f / (mind * mind);
    lt_imp r;
    r.max = visible ? fabsf(minimum_incidence * energy * minimum_outgoing * geometric) : 0.0f;
    const float tpu_c = tc * cos_theta_u - ts * sin_theta_u;
    const int minimum_is_zero = (theta_e - theta_o < 0.0f) || (tc < 0.0f) || (cos_theta_u < 0.0f) ||
                                (tpu_c < cosf(theta_e - theta_o));
    const float tpu_s = lt_sin_from_cos(fclamp(tpu_c, -1.0f, 1.0f));
    const float maximum_outgoing = tpu_c * to_c - tpu_s * to_s;
    r.min = (minimum_is_zero || !visible) ? 0.0f
                                          : fabsf(maximum_incidence * energy * maximum_outgoing / (maxd * maxd));
    return r;
}

/* measure_importance (:185-283), surface variant */
static lt_imp lt_measure_importance(const lt_measure *m, v3 point, v3 n, int has_transmission) {
    lt_imp r = {0.0f, 0.0f};
    if (!(m->has_orientation && m->energy > 0.0f)) return r;
    const v3 cone_axis = lt_safe_normalize(m->axis, V3(0, 0, 1));
    v3 p2c = vneg(cone_axis);
    float cos_theta_u = cosf(m->theta_o + m->theta_e);
    float emitter_distance = 1.0f;
    int visible = 1;
    if (!m->distant) {
        const v3 centroid = vscale(vadd(m->bmin, m->bmax), 0.5f);
        const v3 extent = vsub(m->bmax, centroid);
        const v3 an = V3(fabsf(n.x), fabsf(n.y), fabsf(n.z));
        const v3 ae = V3(fabsf(extent.x), fabsf(extent.y), fabsf(extent.z));
        visible = has_transmission || (vdot(n, vsub(centroid, point)) + vdot(an, ae) > 0.0f);
        const v3 delta = vsub(centroid, point);
        emitter_distance = sqrtf(vdot(delta, delta));
        p2c = lt_safe_normalize(delta, vneg(cone_axis));
        cos_theta_u = lt_cos_bound(m->bmax, centroid, point);
        const v3 cb = vsub(centroid, m->bmax);
        emitter_distance = fmaxf(0.5f * sqrtf(vdot(cb, cb)), emitter_distance);
    }
    if (visible)
        r = lt_bounded(n, has_transmission, p2c, cos_theta_u, cone_axis, m->theta_o, m->theta_e, emitter_distance,
                       emitter_distance, m->energy);
    return r;
}

static lt_imp lt_emitter_importance(const oracle_light_tree *t, uint32_t i, v3 point, v3 n, int has_transmission) {
    const oracle_tree_record *e = &t->emitters[i];
    const lt_measure m = lt_unpack(e->measure, e->flags, e->distant != 0u);
    return lt_measure_importance(&m, point, n, has_transmission);
}

/* child_importance (:317-350) */
static lt_imp lt_child_importance(const oracle_light_tree *t, uint32_t index, v3 point, v3 n, int has_transmission) {
    const oracle_tree_node *node = &t->nodes[index];
    lt_imp r = {0.0f, 0.0f};
    if (node->kind == ORACLE_TREE_INNER || node->count > 1u) {
        const lt_measure m = lt_unpack(node->measure, node->flags, node->kind == ORACLE_TREE_DISTANT);
        r = lt_measure_importance(&m, point, n, has_transmission);
    } else if (node->count == 1u) {
        r = lt_emitter_importance(t, node->first, point, n, has_transmission);
    }
    return r;
}

/* left_probability (:352-386): (probability of the left child, usable) */
static void lt_left_probability(const oracle_light_tree *t, uint32_t left, uint32_t right, v3 point, v3 n,
                                int has_transmission, float *p_left, float *usable) {
    const lt_imp li = lt_child_importance(t, left, point, n, has_transmission);
    const lt_imp ri = lt_child_importance(t, right, point, n, has_transmission);
    const float max_total = li.max + ri.max;
    const float min_total = li.min + ri.min;
    const float max_p = li.max / fmaxf(max_total, LT_SAFE);
    const float min_p = min_total > 0.0f ? li.min / fmaxf(min_total, LT_SAFE)
                                         : 0.5f * ((li.max > 0.0f ? 1.0f : 0.0f) + (ri.max == 0.0f ? 1.0f : 0.0f));
    *p_left = 0.5f * (max_p + min_p);
    *usable = max_total > 0.0f ? 1.0f : 0.0f;
}

static void lt_reservoir(uint32_t current, float weight, uint32_t *selected, float *selected_weight, float *total,
                         float *random) {
    if (!(weight > 0.0f)) return;
    *total += weight;
    if (*selected == LT_INVALID) {
        *selected = current;
        *selected_weight = weight;
        return;
    }
    const float threshold = weight / *total;
    if (*random <= threshold) {
        *selected = current;
        *selected_weight = weight;
        *random = *random / threshold;
    } else {
        *random = (*random - threshold) / fmaxf(1.0f - threshold, LT_SAFE);
    }
    *random = fclamp(*random, 0.0f, 1.0f);
}

/* select_leaf_emitter (:388-535) */
static void lt_select_leaf(const oracle_light_tree *t, const oracle_tree_node *node, v3 point, v3 n,
                           int has_transmission, float *random, uint32_t *selected, float *pdf) {
    float sel_max = 0.0f, sel_min = 0.0f, tot_max = 0.0f, tot_min = 0.0f;
    uint32_t positive = 0u;
    const int sample_maximum = *random > 0.5f;
    if (node->count > 1u) *random = *random * 2.0f - (sample_maximum ? 1.0f : 0.0f);
    for (uint32_t i = 0; i < node->count; ++i) {
        const uint32_t idx = node->first + i;
        const lt_imp imp = lt_emitter_importance(t, idx, point, n, has_transmission);
        tot_max += imp.max;
        tot_min += imp.min;
        positive += imp.max > 0.0f ? 1u : 0u;
        const float primary = sample_maximum ? imp.max : imp.min;
        const float primary_total = sample_maximum ? tot_max : tot_min;
        if (primary > 0.0f) {
            if (*selected == LT_INVALID) {
                *selected = idx;
                sel_max = imp.max;
                sel_min = imp.min;
            } else {
                const float threshold = primary / primary_total;
                if (*random <= threshold) {
                    *selected = idx;
                    sel_max = imp.max;
                    sel_min = imp.min;
                    *random = *random / threshold;
                } else {
                    *random = (*random - threshold) / fmaxf(1.0f - threshold, LT_SAFE);
                }
                *random = fclamp(*random, 0.0f, 1.0f);
            }
        }
    }
    if (positive > 0u && tot_min == 0.0f) {
        if (sample_maximum) {
            sel_min = 1.0f;
            tot_min = (float)positive;
        } else {
            *selected = LT_INVALID;
            float sel_uniform = 0.0f, tot_uniform = 0.0f;
            for (uint32_t i = 0; i < node->count; ++i) {
                const uint32_t idx = node->first + i;
                const lt_imp imp = lt_emitter_importance(t, idx, point, n, has_transmission);
                lt_reservoir(idx, imp.max > 0.0f ? 1.0f : 0.0f, selected, &sel_uniform, &tot_uniform, random);
                if (*selected == idx) {
                    sel_max = imp.max;
                    sel_min = 1.0f;
                }
            }
            tot_min = tot_uniform;
        }
    }
    if (*selected != LT_INVALID)
        *pdf *= 0.5f * (sel_max / fmaxf(tot_max, LT_SAFE) + sel_min / fmaxf(tot_min, LT_SAFE));
}

/* make_sample_callable (:551-637): returns the emitter id (LT_INVALID when
 * nothing is selected) and its selection pdf. */
static uint32_t lt_sample(const oracle_light_tree *t, float u, v3 point, v3 n, int has_transmission, float *pdf) {
    *pdf = 0.0f;
    if (t->node_count == 0u || t->emitter_count == 0u || t->root >= t->node_count) return LT_INVALID;
    float random = fclamp(u, 0.0f, 1.0f);
    float selection_pdf = 1.0f;
    uint32_t index = t->root, selected = LT_INVALID, steps = 0u;
    int active = 1;
    while (active && steps <= t->node_count) {
        steps += 1u;
        const oracle_tree_node *node = &t->nodes[index];
        if (node->kind != ORACLE_TREE_INNER) {
            lt_select_leaf(t, node, point, n, has_transmission, &random, &selected, &selection_pdf);
            active = 0;
        } else {
            float p_left, usable;
            lt_left_probability(t, node->left, node->right, point, n, has_transmission, &p_left, &usable);
            if (usable <= 0.0f) {
                active = 0;
            } else {
                const float p_right = 1.0f - p_left;
                const int choose_right = random <= p_right;
                index = choose_right ? node->right : node->left;
                selection_pdf *= choose_right ? p_right : p_left;
                random = choose_right ? random / fmaxf(p_right, LT_SAFE)
                                      : (random - p_right) / fmaxf(p_left, LT_SAFE);
                random = fclamp(random, 0.0f, 1.0f);
            }
        }
    }
    if (selected == LT_INVALID) return LT_INVALID;
    *pdf = selection_pdf;
    return t->emitters[selected].emitter_id;
}

/* make_pdf_callable (:639-744) */
static float lt_pdf(const oracle_light_tree *t, uint32_t emitter_id, v3 point, v3 n, int has_transmission) {
    if (t->node_count == 0u || t->emitter_count == 0u || emitter_id >= t->emitter_count) return 0.0f;
    const uint32_t target = t->emitter_to_tree[emitter_id];
    uint32_t current = t->emitter_to_leaf[emitter_id];
    if (target >= t->emitter_count || current >= t->node_count) return 0.0f;
    const oracle_tree_node *leaf = &t->nodes[current];
    float tmax = 0.0f, tmin = 0.0f, tot_max = 0.0f, tot_min = 0.0f;
    uint32_t positive = 0u;
    for (uint32_t i = 0; i < leaf->count; ++i) {
        const uint32_t idx = leaf->first + i;
        const lt_imp imp = lt_emitter_importance(t, idx, point, n, has_transmission);
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
                           (tot_min > 0.0f ? tmin / fmaxf(tot_min, LT_SAFE)
                                           : 1.0f / fmaxf((float)positive, 1.0f)));
    uint32_t steps = 0u;
    int valid = 1;
    while (valid && current != t->root && steps <= t->node_count) {
        steps += 1u;
        const uint32_t parent_index = t->nodes[current].parent;
        if (parent_index >= t->node_count) {
            valid = 0;
            result = 0.0f;
        } else {
            const oracle_tree_node *parent = &t->nodes[parent_index];
            float p_left, usable;
            lt_left_probability(t, parent->left, parent->right, point, n, has_transmission, &p_left, &usable);
            if (usable <= 0.0f) {
                valid = 0;
                result = 0.0f;
            } else {
                result *= current == parent->left ? p_left : 1.0f - p_left;
                current = parent_index;
            }
        }
    }
    if (current != t->root) result = 0.0f;
    return result;
}

/* LightDesc -> LightGpu (src/luisa/path_tracer_scene.cpp:1735-1870) */
static void lamp_from_desc(const psx_light *l, olamp *lp) {
    lp->type = l->type;
    float len[3];
    v3 axis[3];
    for (int c = 0; c < 3; ++c) {
        const v3 a = V3(l->transform[4 * c], l->transform[4 * c + 1], l->transform[4 * c + 2]);
        const float n = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
        if (n <= 1.0e-20f) {
            axis[c] = V3(0, 0, 0);
            len[c] = 0.0f;
        } else {
            axis[c] = V3(a.x / n, a.y / n, a.z / n);
            len[c] = n;
        }
    }
    if (axis[2].x == 0.0f && axis[2].y == 0.0f && axis[2].z == 0.0f) axis[2] = V3(0, 0, 1);
    lp->position = V3(l->transform[12], l->transform[13], l->transform[14]);
    lp->axis_x = axis[0];
    lp->axis_y = axis[1];
    lp->axis_z = axis[2];
    lp->axis_scale = V3(len[0], len[1], len[2]);
    lp->color = load3(l->color);
    lp->power = l->power;
    lp->radius = l->size;
    lp->spot_angle = l->spot_angle;
    lp->spot_smooth = l->spot_smooth;
    lp->angle = l->angle;
    lp->visibility = l->visibility_mask;
    lp->max_bounces = l->max_bounces;
    lp->size_u = l->size * len[0];
    lp->size_v = (l->size_y > 0.0f ? l->size_y : l->size) * len[1];
    lp->spread = l->spread;
    const int use_mis = (l->flags & PSX_LIGHT_FLAG_USE_MIS) != 0u;
    int effective_mis;
    if (l->type == PSX_LIGHT_DISTANT)
        effective_mis = use_mis && l->angle > 0.0f;
    else if (l->type == PSX_LIGHT_AREA)
        effective_mis = use_mis && lp->size_u != 0.0f && lp->size_v != 0.0f && l->spread > 0.0f;
    else
        effective_mis = use_mis && l->size > 0.0f;
    lp->flags = ((l->flags & PSX_LIGHT_FLAG_NORMALIZE) ? LAMP_NORMALIZE : 0u) |
                ((l->flags & PSX_LIGHT_FLAG_ELLIPSE) ? LAMP_ELLIPSE : 0u) |
                ((l->flags & PSX_LIGHT_FLAG_SPHERE) ? LAMP_SPHERE : 0u) |
                (effective_mis ? LAMP_USE_MIS : 0u) |
                (effective_mis && l->type != PSX_LIGHT_DISTANT ? LAMP_FORWARD : 0u) |
                (l->type == PSX_LIGHT_AREA && l->spread >= 3.14159265358979f - 1.0e-6f ? LAMP_FULL_SPREAD : 0u) |
                ((l->flags & PSX_LIGHT_FLAG_CAST_SHADOW) ? LAMP_CAST_SHADOW : 0u);
}

oracle_scene *oracle_scene_create(const psx_scene_desc *d, int brute_force) {
    if (!d) return NULL;
    oracle_scene *s = (oracle_scene *)calloc(1, sizeof(oracle_scene));
    s->brute_force = brute_force;
    s->camera = d->camera;
    s->world_visibility = d->world_visibility_mask;
    s->background = load3(d->world_color);
    /* background lights fold into the constant world radiance; point, spot
     * and distant lamps become analytic emitters
     * (src/luisa/path_tracer_scene.cpp:1723-1870) */
    s->lamps = (olamp *)calloc(d->light_count ? d->light_count : 1, sizeof(olamp));
    for (uint32_t i = 0; i < d->light_count; ++i) {
        const psx_light *l = &d->lights[i];
        if (l->type == PSX_LIGHT_BACKGROUND) {
            s->background.x += l->color[0] * l->power;
            s->background.y += l->color[1] * l->power;
            s->background.z += l->color[2] * l->power;
            continue;
        }
        if (l->type > PSX_LIGHT_DISTANT) continue;
        lamp_from_desc(l, &s->lamps[s->lamp_count++]);
    }
    s->material_count = d->material_count;
    s->materials = (omaterial *)calloc(d->material_count ? d->material_count : 1, sizeof(omaterial));
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const psx_material *src = &d->materials[i];
        omaterial *m = &s->materials[i];
        m->count = src->instruction_count;
        m->root = src->root;
        m->instr = (psx_closure_instr *)dupmem(src->instructions,
                                               sizeof(psx_closure_instr) * src->instruction_count);
        m->param_count = src->parameter_count;
        m->params = (float *)dupmem(src->parameters, sizeof(float) * src->parameter_count);
        m->emission_sampling = src->emission_sampling;
        material_refresh(m);
    }
    s->mesh_count = d->mesh_count;
    s->meshes = (omesh *)calloc(d->mesh_count ? d->mesh_count : 1, sizeof(omesh));
    uint32_t prim_offset = 0;
    for (uint32_t i = 0; i < d->mesh_count; ++i) {
        const psx_mesh *src = &d->meshes[i];
        omesh *m = &s->meshes[i];
        m->vcount = src->vertex_count;
        m->tcount = src->triangle_count;
        m->positions = (float *)dupmem(src->positions, sizeof(float) * 3u * src->vertex_count);
        m->normals = (float *)dupmem(src->normals, sizeof(float) * 3u * src->vertex_count);
        m->tris = (uint32_t *)dupmem(src->triangles, sizeof(uint32_t) * 3u * src->triangle_count);
        m->smooth = (uint8_t *)dupmem(src->triangle_smooth, src->triangle_count);
        m->slots = (uint32_t *)dupmem(src->material_slots, sizeof(uint32_t) * src->material_slot_count);
        m->slot_count = src->material_slot_count;
        m->tri_slots = (uint32_t *)dupmem(src->triangle_material_slots,
                                          sizeof(uint32_t) * src->triangle_count);
        m->prim_offset = prim_offset;
        prim_offset += src->triangle_count;
    }
    for (uint32_t i = 0; i < d->instance_count; ++i)
        if (d->instances[i].mesh < d->mesh_count) s->meshes[d->instances[i].mesh].users++;

    /* BLASes: one object-space BLAS per shared mesh, one world-space BLAS per
     * single-user mesh (build_cycles_instance_intersection_plan,
     * src/luisa/path_tracer_scene_geometry.cpp:272-318). */
    s->blases = (blas *)calloc(d->instance_count + d->mesh_count + 1, sizeof(blas));
    uint32_t *shared_blas = (uint32_t *)malloc(sizeof(uint32_t) * (d->mesh_count + 1));
    for (uint32_t i = 0; i < d->mesh_count; ++i) shared_blas[i] = UINT32_MAX;
    s->instance_count = d->instance_count;
    s->instances = (oinstance *)calloc(d->instance_count ? d->instance_count : 1, sizeof(oinstance));
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const psx_instance *src = &d->instances[i];
        oinstance *inst = &s->instances[i];
        inst->mesh = src->mesh;
        inst->visibility = src->visibility_mask;
        memcpy(inst->o2w.m, src->transform, sizeof(inst->o2w.m));
        inst->w2o = cycles_inverse(&inst->o2w);
        inst->n2w = transpose4(&inst->w2o);
        inst->override = (uint32_t *)dupmem(src->material_overrides,
                                            sizeof(uint32_t) * src->material_override_count);
        inst->override_count = src->material_override_count;
        inst->terminator_offset = src->shadow_terminator_geometry_offset;
        const omesh *m = &s->meshes[src->mesh];
        inst->applied = m->users == 1u;
        if (inst->applied) {
            blas *b = &s->blases[s->blas_count];
            b->mesh = src->mesh;
            b->world = 1;
            b->ipos = (float *)malloc(sizeof(float) * 3u * (m->vcount ? m->vcount : 1));
            for (uint32_t v = 0; v < m->vcount; ++v) {
                v3 p = xform_point(&inst->o2w, load3(m->positions + 3u * v));
                b->ipos[3u * v] = p.x;
                b->ipos[3u * v + 1] = p.y;
                b->ipos[3u * v + 2] = p.z;
            }
            build_blas(b, m);
            inst->blas = s->blas_count++;
        } else {
            if (shared_blas[src->mesh] == UINT32_MAX) {
                blas *b = &s->blases[s->blas_count];
                b->mesh = src->mesh;
                b->world = 0;
                b->ipos = (float *)dupmem(m->positions, sizeof(float) * 3u * m->vcount);
                build_blas(b, m);
                shared_blas[src->mesh] = s->blas_count++;
            }
            inst->blas = shared_blas[src->mesh];
        }
        /* world bounds of the instance */
        const blas *b = &s->blases[inst->blas];
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        if (b->root_lo[0] <= b->root_hi[0]) {
            for (int c = 0; c < 8; ++c) {
                v3 corner = V3((c & 1) ? b->root_hi[0] : b->root_lo[0],
                               (c & 2) ? b->root_hi[1] : b->root_lo[1],
                               (c & 4) ? b->root_hi[2] : b->root_lo[2]);
                v3 w = b->world ? corner : xform_point(&inst->o2w, corner);
                lo[0] = fminf(lo[0], w.x); lo[1] = fminf(lo[1], w.y); lo[2] = fminf(lo[2], w.z);
                hi[0] = fmaxf(hi[0], w.x); hi[1] = fmaxf(hi[1], w.y); hi[2] = fmaxf(hi[2], w.z);
            }
            pad_box(lo, hi);
        }
        memcpy(inst->lo, lo, sizeof(lo));
        memcpy(inst->hi, hi, sizeof(hi));
    }
    free(shared_blas);
    /* TLAS */
    {
        build_item *items = (build_item *)malloc(sizeof(build_item) * (d->instance_count ? d->instance_count : 1));
        uint32_t n = 0;
        for (uint32_t i = 0; i < d->instance_count; ++i) {
            if (s->instances[i].lo[0] > s->instances[i].hi[0]) continue;
            memcpy(items[n].lo, s->instances[i].lo, sizeof(items[n].lo));
            memcpy(items[n].hi, s->instances[i].hi, sizeof(items[n].hi));
            for (int a = 0; a < 3; ++a) items[n].c[a] = 0.5f * (items[n].lo[a] + items[n].hi[a]);
            items[n].id = i;
            ++n;
        }
        s->tlas_prims = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
        build_bvh(items, n, &s->tlas, &s->tlas_count, s->tlas_prims);
        free(items);
    }
    /* emissive triangles (src/luisa/path_tracer_scene.cpp:1548-1640) */
    {
        uint32_t cap = 16, n = 0;
        emitter *em = (emitter *)malloc(sizeof(emitter) * cap);
        float *areas = (float *)malloc(sizeof(float) * cap);
        for (uint32_t i = 0; i < d->instance_count; ++i) {
            const oinstance *inst = &s->instances[i];
            const int light_visible = inst->visibility == UINT32_MAX ||
                                      (inst->visibility & (PSX_VISIBILITY_DIFFUSE | PSX_VISIBILITY_GLOSSY |
                                                           PSX_VISIBILITY_TRANSMISSION |
                                                           PSX_VISIBILITY_VOLUME_SCATTER)) != 0u;
            if (!light_visible) continue;
            const omesh *m = &s->meshes[inst->mesh];
            for (uint32_t p = 0; p < m->tcount; ++p) {
                const uint32_t mat = material_of(s, inst, p);
                if (mat >= s->material_count || !s->materials[mat].may_emit) continue;
                if (n == cap) {
                    cap *= 2;
                    em = (emitter *)realloc(em, sizeof(emitter) * cap);
                    areas = (float *)realloc(areas, sizeof(float) * cap);
                }
                em[n].instance = i;
                em[n].prim = p;
                em[n].material = mat;
                em[n].emission_sampling = s->materials[mat].emission_sampling;
                v3 p0, p1, p2;
                tri_positions(m->positions, m->tris, p, &p0, &p1, &p2);
                p0 = xform_point(&inst->o2w, p0);
                p1 = xform_point(&inst->o2w, p1);
                p2 = xform_point(&inst->o2w, p2);
                const v3 n01 = vsub(p1, p0), n02 = vsub(p2, p0);
                const v3 nn = vcross(n01, n02);
                areas[n] = 0.5f * sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
                ++n;
            }
        }
        s->emitters = em;
        s->emitter_count = n;
        s->distribution = (dist_entry *)malloc(sizeof(dist_entry) * (n + s->lamp_count + 2u));
        s->distribution_count = build_distribution(areas, n, s->lamp_count, 0, s->distribution,
                                                   &s->triangle_area_pdf,
                                                   &s->light_selection_pdf);
        if (!(s->distribution[s->distribution_count].cumulative > 0.0f))
            s->distribution_count = 0u;
        free(areas);
    }
    /* light tree over the same population: triangles, then lamps
     * (src/luisa/path_tracer_light_sampling_scene.cpp:150-245) */
    {
        const uint32_t total = s->emitter_count + s->lamp_count;
        oracle_tree_emitter *te = (oracle_tree_emitter *)calloc(total ? total : 1, sizeof(oracle_tree_emitter));
        s->emitter_cprim = (uint32_t *)calloc(s->emitter_count ? s->emitter_count : 1, sizeof(uint32_t));
        for (uint32_t i = 0; i < s->emitter_count; ++i) {
            const emitter *em = &s->emitters[i];
            const oinstance *inst = &s->instances[em->instance];
            const omesh *m = &s->meshes[inst->mesh];
            v3 p0, p1, p2;
            tri_positions(m->positions, m->tris, em->prim, &p0, &p1, &p2);
            const omaterial *mat = &s->materials[em->material];
            const v3 estimate = mat->count ? emission_estimate(mat, mat->root) : V3(0, 0, 0);
            te[i] = triangle_tree_emitter(i, inst, p0, p1, p2, estimate, em->emission_sampling);
            s->emitter_cprim[i] = m->prim_offset + em->prim;
        }
        for (uint32_t j = 0; j < s->lamp_count; ++j)
            te[s->emitter_count + j] = lamp_tree_emitter(s->emitter_count + j, &s->lamps[j]);
        if (oracle_light_tree_build(te, total, 8u, &s->tree) != 0) memset(&s->tree, 0, sizeof(s->tree));
        if (!(s->tree.node_count > 0u && s->tree.nodes[s->tree.root].measure[11] > 0.0f)) {
            oracle_light_tree_free(&s->tree);
        }
        free(te);
    }
    return s;
}

void oracle_scene_destroy(oracle_scene *s) {
    if (!s) return;
    for (uint32_t i = 0; i < s->material_count; ++i) {
        free(s->materials[i].instr);
        free(s->materials[i].params);
    }
    free(s->materials);
    for (uint32_t i = 0; i < s->mesh_count; ++i) {
        omesh *m = &s->meshes[i];
        free(m->positions); free(m->normals); free(m->tris); free(m->smooth);
        free(m->slots); free(m->tri_slots);
    }
    free(s->meshes);
    for (uint32_t i = 0; i < s->instance_count; ++i) free(s->instances[i].override);
    free(s->instances);
    for (uint32_t i = 0; i < s->blas_count; ++i) {
        free(s->blases[i].ipos); free(s->blases[i].nodes); free(s->blases[i].prims);
    }
    free(s->blases);
    free(s->tlas);
    free(s->tlas_prims);
    free(s->emitters);
    free(s->distribution);
    free(s->lamps);
    free(s->emitter_cprim);
    oracle_light_tree_free(&s->tree);
    free(s);
}

int oracle_scene_update_parameters(oracle_scene *s, uint32_t material,
                                   const float *parameters, uint32_t count) {
    if (!s || material >= s->material_count) return -1;
    omaterial *m = &s->materials[material];
    if (count != m->param_count) return -1;
    memcpy(m->params, parameters, sizeof(float) * count);
    material_refresh(m);
    return 0;
}

uint32_t oracle_scene_emissive_triangles(const oracle_scene *s) {
    return s ? s->emitter_count : 0u;
}

uint32_t oracle_scene_distribution(const oracle_scene *s, float *cumulative, float *selection_pdf, uint32_t *kind,
                                   uint32_t *index, uint32_t capacity) {
    if (!s) return 0u;
    const uint32_t n = s->distribution_count;
    for (uint32_t i = 0; i < n && i < capacity; ++i) {
        cumulative[i] = s->distribution[i].cumulative;
        selection_pdf[i] = s->distribution[i].selection_pdf;
        kind[i] = s->distribution[i].kind;
        index[i] = s->distribution[i].index;
    }
    return n;
}

int oracle_scene_light_tree(const oracle_scene *s, oracle_light_tree *view) {
    if (!s || !view) return -1;
    *view = s->tree;
    return s->tree.node_count > 0u ? 0 : 1;
}

/* ======================================================================= */
/* closest hit (src/luisa/path_kernel_scene_traversal.cpp:34-68, 212-493)  */
/* ======================================================================= */

typedef struct ray_t {
    v3 o, d;
    float tmin, tmax;
} ray_t;

typedef struct hit_t {
    int valid;
    uint32_t inst, prim;  /* prim is local to the mesh */
    float t, u, v;
    uint32_t object, cprim;
} hit_t;

typedef struct trace_counters {
    uint64_t nodes, tris;
} trace_counters;

typedef struct closest_state {
    hit_t best;
    uint32_t visibility;
    uint32_t src_obj, src_prim, light_obj, light_prim;
    float tmin, tmax;  /* original closed interval */
    trace_counters *counters;
} closest_state;

/* CyclesClosestHitOrder::accepts */
static int order_accepts(const hit_t *best, float t, uint32_t object, uint32_t prim) {
    if (!best->valid) return 1;
    if (t < best->t) return 1;
    if (t == best->t) {
        if (object > best->object) return 1;
        if (object == best->object && prim > best->cprim) return 1;
    }
    return 0;
}

static void test_triangle(const oracle_scene *s, closest_state *st, uint32_t inst_index,
                          const oinstance *inst, const blas *b, v3 o, v3 d, uint32_t prim) {
    const omesh *m = &s->meshes[inst->mesh];
    if (st->counters) st->counters->tris++;
    v3 p0, p1, p2;
    tri_positions(b->ipos, m->tris, prim, &p0, &p1, &p2);
    const tri_hit h = pluecker(o, d, st->tmin, st->tmax, p0, p1, p2);
    if (!h.valid) return;
    const uint32_t object = inst_index;
    const uint32_t cprim = m->prim_offset + prim;
    const int excluded = (object == st->src_obj && cprim == st->src_prim) ||
                         (object == st->light_obj && cprim == st->light_prim);
    if (excluded) return;
    if ((inst->visibility & st->visibility) == 0u) return;
    if (!order_accepts(&st->best, h.t, object, cprim)) return;
    st->best.valid = 1;
    st->best.inst = inst_index;
    st->best.prim = prim;
    st->best.t = h.t;
    st->best.u = h.u;
    st->best.v = h.v;
    st->best.object = object;
    st->best.cprim = cprim;
}

/* conservative slab test: the box may only be culled when no triangle in it
 * can satisfy the exact test within [tmin, best]. */
static int box_hit(const float lo[3], const float hi[3], v3 o, v3 inv, float tmin, float tmax) {
    float t0x = (lo[0] - o.x) * inv.x, t1x = (hi[0] - o.x) * inv.x;
    float t0y = (lo[1] - o.y) * inv.y, t1y = (hi[1] - o.y) * inv.y;
    float t0z = (lo[2] - o.z) * inv.z, t1z = (hi[2] - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    const float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    const float tf_safe = tf * 1.00001f + 1.0e-30f;
    return tn <= tf_safe && tf_safe >= tmin && tn <= tmax;
}

static float query_upper(float t) {
    /* widen the pruning bound so that ties and rounding stay visible */
    return t * 1.00001f + 1.0e-30f;
}

static void traverse_blas(const oracle_scene *s, closest_state *st, uint32_t inst_index,
                          const oinstance *inst, v3 o, v3 d) {
    const blas *b = &s->blases[inst->blas];
    const omesh *m = &s->meshes[inst->mesh];
    if (s->brute_force) {
        for (uint32_t p = 0; p < m->tcount; ++p) test_triangle(s, st, inst_index, inst, b, o, d, p);
        return;
    }
    const v3 dc = V3(clamp_bvh_direction(d.x), clamp_bvh_direction(d.y), clamp_bvh_direction(d.z));
    const v3 inv = V3(1.0f / dc.x, 1.0f / dc.y, 1.0f / dc.z);
    const float lower = st->tmin > 0.0f ? st->tmin * 0.99999f : st->tmin - 1.0e-30f;
    uint32_t stack[128];
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const uint32_t ni = stack[--sp];
        const bvh_node *n = &b->nodes[ni];
        if (st->counters) st->counters->nodes++;
        const float upper = st->best.valid ? query_upper(st->best.t) : query_upper(st->tmax);
        if (!box_hit(n->lo, n->hi, o, inv, lower, upper)) continue;
        if (n->count) {
            for (uint32_t i = 0; i < n->count; ++i)
                test_triangle(s, st, inst_index, inst, b, o, d, b->prims[n->first + i]);
        } else {
            /* visit the nearer child first */
            const bvh_node *l = &b->nodes[n->first], *r = &b->nodes[n->first + 1];
            const float cl = (l->lo[0] + l->hi[0] - 2.0f * o.x) * dc.x +
                             (l->lo[1] + l->hi[1] - 2.0f * o.y) * dc.y +
                             (l->lo[2] + l->hi[2] - 2.0f * o.z) * dc.z;
            const float cr = (r->lo[0] + r->hi[0] - 2.0f * o.x) * dc.x +
                             (r->lo[1] + r->hi[1] - 2.0f * o.y) * dc.y +
                             (r->lo[2] + r->hi[2] - 2.0f * o.z) * dc.z;
            if (sp + 2 > 128) continue; /* cannot happen for sane depths */
            if (cl <= cr) { stack[sp++] = n->first + 1; stack[sp++] = n->first; }
            else { stack[sp++] = n->first; stack[sp++] = n->first + 1; }
        }
    }
}

static void visit_instance(const oracle_scene *s, closest_state *st, const ray_t *r,
                           uint32_t i) {
    const oinstance *inst = &s->instances[i];
    if ((inst->visibility & st->visibility) == 0u) return;
    v3 o = r->o, d = r->d;
    if (!inst->applied) {
        o = xform_point(&inst->w2o, r->o);
        d = xform_dir(&inst->w2o, r->d);
    }
    traverse_blas(s, st, i, inst, o, d);
}

static hit_t closest_hit(const oracle_scene *s, const ray_t *r, uint32_t visibility,
                         uint32_t src_obj, uint32_t src_prim, uint32_t light_obj,
                         uint32_t light_prim, trace_counters *counters) {
    closest_state st;
    memset(&st, 0, sizeof(st));
    st.visibility = visibility;
    st.src_obj = src_obj;
    st.src_prim = src_prim;
    st.light_obj = light_obj;
    st.light_prim = light_prim;
    st.tmin = r->tmin;
    st.tmax = r->tmax;
    st.counters = counters;
    if (s->brute_force || s->tlas_count == 0) {
        for (uint32_t i = 0; i < s->instance_count; ++i) visit_instance(s, &st, r, i);
        return st.best;
    }
    const v3 dc = V3(clamp_bvh_direction(r->d.x), clamp_bvh_direction(r->d.y),
                     clamp_bvh_direction(r->d.z));
    const v3 inv = V3(1.0f / dc.x, 1.0f / dc.y, 1.0f / dc.z);
    const float lower = r->tmin > 0.0f ? r->tmin * 0.99999f : r->tmin - 1.0e-30f;
    uint32_t stack[128];
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const bvh_node *n = &s->tlas[stack[--sp]];
        if (counters) counters->nodes++;
        const float upper = st.best.valid ? query_upper(st.best.t) : query_upper(r->tmax);
        if (!box_hit(n->lo, n->hi, r->o, inv, lower, upper)) continue;
        if (n->count) {
            for (uint32_t k = 0; k < n->count; ++k) visit_instance(s, &st, r, s->tlas_prims[n->first + k]);
        } else {
            if (sp + 2 > 128) continue;
            stack[sp++] = n->first + 1;
            stack[sp++] = n->first;
        }
    }
    return st.best;
}

/* ======================================================================= */
/* surface reconstruction                                                   */
/* (src/luisa/cycles_triangle_surface_component.cpp:13-103,                */
/*  src/luisa/path_kernel_surface_primitive.cpp:139-460)                   */
/* ======================================================================= */

typedef struct surface_t {
    uint32_t inst, prim, object, cprim, material;
    int applied, smooth, back_facing;
    v3 p0, p1, p2;        /* intersection-space positions (cycles p) */
    v3 cn0, cn1, cn2;     /* intersection-space vertex normals (cycles n) */
    v3 wp0, wp1, wp2;     /* world positions */
    v3 position, object_hit_position;
    v3 ng, ns;
    v3 incoming;
    float u, v;
    uint32_t emission_sampling; /* effective triangle emission sampling */
} surface_t;

static v3 interp3(float u, float v, v3 a, v3 b, v3 c) {
    /* triangle_interpolate: (1 - u - v) a + u b + v c */
    return vadd(vadd(vscale(a, 1.0f - u - v), vscale(b, u)), vscale(c, v));
}

static void resolve_surface(const oracle_scene *s, const hit_t *h, const ray_t *r,
                            surface_t *sf) {
    const oinstance *inst = &s->instances[h->inst];
    const omesh *m = &s->meshes[inst->mesh];
    const blas *b = &s->blases[inst->blas];
    sf->inst = h->inst;
    sf->prim = h->prim;
    sf->object = h->object;
    sf->cprim = h->cprim;
    sf->material = material_of(s, inst, h->prim);
    sf->applied = inst->applied;
    sf->smooth = (m->smooth && m->normals) ? (m->smooth[h->prim] != 0) : 0;
    sf->u = h->u;
    sf->v = h->v;
    v3 op0, op1, op2, fp0, fp1, fp2;
    tri_positions(m->positions, m->tris, h->prim, &op0, &op1, &op2);
    tri_positions(b->ipos, m->tris, h->prim, &fp0, &fp1, &fp2);
    const v3 p0 = sf->applied ? fp0 : op0;
    const v3 p1 = sf->applied ? fp1 : op1;
    const v3 p2 = sf->applied ? fp2 : op2;
    sf->p0 = p0; sf->p1 = p1; sf->p2 = p2;
    const v3 tp0 = xform_point(&inst->o2w, op0);
    const v3 tp1 = xform_point(&inst->o2w, op1);
    const v3 tp2 = xform_point(&inst->o2w, op2);
    sf->wp0 = sf->applied ? p0 : tp0;
    sf->wp1 = sf->applied ? p1 : tp1;
    sf->wp2 = sf->applied ? p2 : tp2;
    const v3 ogn = safe_normalize(vcross(vsub(op1, op0), vsub(op2, op0)), V3(0, 0, 1));
    const v3 tgn = safe_normalize(xform_dir(&inst->n2w, ogn), vneg(r->d));
    v3 sgn = safe_normalize(vcross(vsub(p1, p0), vsub(p2, p0)), tgn);
    if (vdot(sgn, tgn) < 0.0f) sgn = vneg(sgn);
    v3 gn = sf->applied ? sgn : tgn;
    v3 n0 = V3(0, 0, 0), n1 = n0, n2 = n0;
    if (m->normals) {
        const uint32_t *t = m->tris + 3u * h->prim;
        n0 = load3(m->normals + 3u * t[0]);
        n1 = load3(m->normals + 3u * t[1]);
        n2 = load3(m->normals + 3u * t[2]);
    }
    const v3 wn0 = safe_normalize(xform_dir(&inst->n2w, n0), gn);
    const v3 wn1 = safe_normalize(xform_dir(&inst->n2w, n1), gn);
    const v3 wn2 = safe_normalize(xform_dir(&inst->n2w, n2), gn);
    const v3 cn0 = sf->applied ? wn0 : n0;
    const v3 cn1 = sf->applied ? wn1 : n1;
    const v3 cn2 = sf->applied ? wn2 : n2;
    sf->cn0 = cn0; sf->cn1 = cn1; sf->cn2 = cn2;
    const v3 osn = sf->smooth ? interp3(h->u, h->v, n0, n1, n2) : ogn;
    const v3 tsn = safe_normalize(xform_dir(&inst->n2w, osn), gn);
    const v3 ssn = safe_normalize(sf->smooth ? interp3(h->u, h->v, cn0, cn1, cn2) : gn, gn);
    v3 ns = sf->applied ? ssn : tsn;
    const int back_facing = vdot(gn, vneg(r->d)) < 0.0f;
    if (back_facing) { gn = vneg(gn); ns = vneg(ns); }
    if (vdot(ns, gn) < 0.0f) ns = vneg(ns);
    sf->back_facing = back_facing;
    sf->ng = gn;
    sf->ns = ns;
    const v3 object_position =
        vadd(vadd(op0, vscale(vsub(op1, op0), h->u)), vscale(vsub(op2, op0), h->v));
    const v3 support_position =
        vadd(vadd(p0, vscale(vsub(p1, p0), h->u)), vscale(vsub(p2, p0), h->v));
    sf->position = sf->applied ? support_position : xform_point(&inst->o2w, object_position);
    sf->object_hit_position = sf->applied ? sf->position : mat_mul(&inst->w2o, sf->position, 1.0f);
    sf->incoming = vneg(r->d);
    /* triangle emission sampling (src/luisa/path_kernel_primitive_material.cpp:44-66) */
    sf->emission_sampling = PSX_EMISSION_SAMPLING_NONE;
    if (sf->material < s->material_count) {
        const omaterial *mat = &s->materials[sf->material];
        const int light_visible = (inst->visibility & (PSX_VISIBILITY_DIFFUSE | PSX_VISIBILITY_GLOSSY |
                                                       PSX_VISIBILITY_TRANSMISSION |
                                                       PSX_VISIBILITY_VOLUME_SCATTER)) != 0u;
        if (mat->may_emit && light_visible) sf->emission_sampling = mat->emission_sampling;
    }
}

/* offset_ray_origin: Waechter & Binder, "A Fast and Robust Method for
 * Avoiding Self-Intersection", Ray Tracing Gems (2019), chapter 6 - the
 * method behind LuisaCompute's offset_ray_origin used at
 * include/psycles/luisa/surface_ray.h:101. */
static float offset_component(float p, float n) {
    const float origin = 1.0f / 32.0f;
    const float float_scale = 1.0f / 65536.0f;
    const float int_scale = 256.0f;
    const int32_t of_i = (int32_t)(int_scale * n);
    int32_t bits;
    memcpy(&bits, &p, 4);
    bits += p < 0.0f ? -of_i : of_i;
    float p_i;
    memcpy(&p_i, &bits, 4);
    return fabsf(p) < origin ? p + float_scale * n : p_i;
}

static v3 offset_ray_origin(v3 p, v3 n) {
    return V3(offset_component(p.x, n.x), offset_component(p.y, n.y),
              offset_component(p.z, n.z));
}

/* include/psycles/luisa/surface_ray.h:50-106 */
static int source_covers_origin(v3 o, v3 d, v3 p0, v3 p1, v3 p2) {
    const v3 v0 = vsub(p0, o), v1 = vsub(p1, o), v2 = vsub(p2, o);
    const v3 e0 = vsub(v2, v0), e1 = vsub(v0, v1), e2 = vsub(v1, v2);
    const float u = vdot(vcross(vadd(v2, v0), e0), d);
    const float v = vdot(vcross(vadd(v0, v1), e1), d);
    const float w = vdot(vcross(vadd(v1, v2), e2), d);
    const float eps = FLT_EPSILON * fabsf(u + v + w);
    const float mn = fminf(u, fminf(v, w));
    const float mx = fmaxf(u, fmaxf(v, w));
    return (mn >= eps) || (mx <= -eps);
}

static v3 origin_with_self_exclusion(v3 world_origin, v3 world_ng, v3 obj_origin,
                                     v3 obj_dir, v3 p0, v3 p1, v3 p2) {
    return source_covers_origin(obj_origin, obj_dir, p0, p1, p2)
               ? world_origin
               : offset_ray_origin(world_origin, world_ng);
}

/* make_ray_origin (src/luisa/path_kernel_surface_geometry.cpp:151-165) */
static v3 surface_ray_origin(const oracle_scene *s, const surface_t *sf, v3 dir) {
    const oinstance *inst = &s->instances[sf->inst];
    const v3 object_dir = xform_dir(&inst->w2o, dir);
    const v3 co = sf->applied ? sf->position : sf->object_hit_position;
    const v3 cd = sf->applied ? dir : object_dir;
    return origin_with_self_exclusion(sf->position, sf->ng, co, cd, sf->p0, sf->p1, sf->p2);
}

/* Cycles' parabolic smooth-surface height (smooth_surface_offset,
 * include/psycles/luisa/surface_ray.h:108-181): positions and normals in
 * intersection space, the interpolated normal taken to world space with the
 * instance's FMA affine unless the transform is applied. */
static v3 smooth_surface_offset(const mat4 *object_to_world, int applied, float bu, float bv, v3 p0, v3 p1,
                                v3 p2, v3 n0, v3 n1, v3 n2, v3 world_ng) {
    const float u = 1.0f - bu - bv;
    const float v = bu;
    const float w = bv;
    const v3 object_position = vadd(vadd(vscale(p0, u), vscale(p1, v)), vscale(p2, w));
    const v3 object_normal = vadd(vadd(vscale(n0, u), vscale(n1, v)), vscale(n2, w));
    const v3 world_normal = applied ? object_normal : xform_dir(object_to_world, object_normal);
    const float a = vdot(vsub(n2, n0), vsub(p0, p2));
    const float b = vdot(vsub(n2, n1), vsub(p1, p2));
    const float c = vdot(vsub(n1, n0), vsub(p1, p0));
    const float parabolic = ((a * u) * (u - 1.0f) + ((a + b) + c) * u * v) + (b * v) * (v - 1.0f);
    float f0 = fmaxf(fmaxf(vdot(vsub(p1, p0), n0), vdot(vsub(p2, p0), n0)), 0.0f);
    float f1 = fmaxf(fmaxf(vdot(vsub(p0, p1), n1), vdot(vsub(p2, p1), n1)), 0.0f);
    float f2 = fmaxf(fmaxf(vdot(vsub(p0, p2), n2), vdot(vsub(p1, p2), n2)), 0.0f);
    f0 = fmaxf(vdot(vsub(p0, object_position), n0) + f0, 0.0f);
    f1 = fmaxf(vdot(vsub(p1, object_position), n1) + f1, 0.0f);
    f2 = fmaxf(vdot(vsub(p2, object_position), n2) + f2, 0.0f);
    const float front = fmaxf(fminf(f0, fminf(f1, f2)), parabolic * 0.5f);
    float b0 = fmaxf(fmaxf(vdot(vsub(p0, p1), n0), vdot(vsub(p0, p2), n0)), 0.0f);
    float b1 = fmaxf(fmaxf(vdot(vsub(p1, p0), n1), vdot(vsub(p1, p2), n1)), 0.0f);
    float b2 = fmaxf(fmaxf(vdot(vsub(p2, p0), n2), vdot(vsub(p2, p1), n2)), 0.0f);
    b0 = fmaxf(vdot(vsub(object_position, p0), n0) + b0, 0.0f);
    b1 = fmaxf(vdot(vsub(object_position, p1), n1) + b1, 0.0f);
    b2 = fmaxf(vdot(vsub(object_position, p2), n2) + b2, 0.0f);
    const float back = fminf(-fminf(b0, fminf(b1, b2)), parabolic * 0.5f);
    const float height = vdot(world_normal, world_ng) > 0.0f ? front : back;
    return vscale(world_normal, height);
}

typedef struct shadow_origin_t {
    v3 position;
    int skip_self;
} shadow_origin_t;

/* shadow_terminator_origin (include/psycles/luisa/surface_ray.h:183-243):
 * Cycles' shadow-terminator geometry offset; the source primitive stays
 * excluded unless the lifted origin faces away from the light. */
static shadow_origin_t shadow_terminator_origin(v3 world_position, v3 world_ns, v3 world_ng, v3 light_dir,
                                                float geometry_offset, int smooth, const mat4 *o2w, int applied,
                                                float bu, float bv, v3 p0, v3 p1, v3 p2, v3 n0, v3 n1, v3 n2) {
    float normal_light = vdot(world_ns, light_dir);
    const int transmit = normal_light < 0.0f;
    normal_light = fabsf(normal_light);
    const v3 offset_normal = transmit ? vneg(world_ng) : world_ng;
    const float geometric_light = vdot(offset_normal, light_dir);
    const float safe_cutoff = fmaxf(geometry_offset, 1.0e-20f);
    const float near_terminator = fclamp(2.0f - (geometric_light + normal_light) / safe_cutoff, 0.0f, 1.0f);
    const float regular = fclamp(1.0f - geometric_light / safe_cutoff, 0.0f, 1.0f);
    const float amount = normal_light < geometry_offset ? near_terminator : regular;
    const int active = smooth && geometry_offset > 0.0f && amount > 0.0f;
    shadow_origin_t r;
    r.position = world_position;
    r.skip_self = 1;
    if (active) {
        const v3 offset = smooth_surface_offset(o2w, applied, bu, bv, p0, p1, p2, n0, n1, n2, offset_normal);
        r.position = vadd(world_position, vscale(offset, amount));
        r.skip_self = geometric_light > 0.0f;
    }
    return r;
}

/* make_shadow_origin (src/luisa/path_kernel_surface_geometry.cpp:168-182)
 * and surface_shadow_origin (include/psycles/luisa/surface_ray.h:252-309):
 * the terminator lift, then the source-triangle certificate with the robust
 * offset fallback while the source primitive stays excluded. */
static shadow_origin_t shadow_origin_core(v3 position, v3 ns, v3 ng, v3 dir, float geometry_offset, int smooth,
                                          const mat4 *o2w, const mat4 *w2o, int applied, float bu, float bv,
                                          v3 p0, v3 p1, v3 p2, v3 n0, v3 n1, v3 n2) {
    shadow_origin_t o = shadow_terminator_origin(position, ns, ng, dir, geometry_offset, smooth, o2w, applied, bu,
                                                 bv, p0, p1, p2, n0, n1, n2);
    const v3 oo = applied ? o.position : xform_point(w2o, o.position);
    const v3 od = applied ? dir : xform_dir(w2o, dir);
    const v3 offset = origin_with_self_exclusion(o.position, ng, oo, od, p0, p1, p2);
    if (o.skip_self) o.position = offset;
    return o;
}

static shadow_origin_t surface_shadow_origin(const oracle_scene *s, const surface_t *sf, v3 dir) {
    const oinstance *inst = &s->instances[sf->inst];
    return shadow_origin_core(sf->position, sf->ns, sf->ng, dir, inst->terminator_offset, sf->smooth, &inst->o2w,
                              &inst->w2o, sf->applied, sf->u, sf->v, sf->p0, sf->p1, sf->p2, sf->cn0, sf->cn1,
                              sf->cn2);
}

static uint32_t intersection_t_offset_bits(float t) {
    return t == 0.0f ? fbits(FLT_MIN) : fbits(t) + 1u;
}

/* ======================================================================= */
/* closure evaluation and sampling                                          */
/* (src/luisa/surface_closure_evaluation.cpp:67-464,                       */
/*  src/luisa/surface_closure_sampling.cpp:11-95,                          */
/*  src/luisa/surface_closure_sampling_core.cpp:96-676,                    */
/*  src/luisa/graph_surface_scattering.cpp:463-506)                        */
/* ======================================================================= */

static float oren_nayar_g(float cosine) {
    const float c = fclamp(cosine, 0.0f, 1.0f);
    const float sine = sqrtf(fmaxf(1.0f - c * c, 0.0f));
    const float theta = acosf(c);
    const float safe_c = fmaxf(c, 1.0e-6f);
    const float regular = sine * (theta - 2.0f / 3.0f - sine * c) +
                          2.0f / 3.0f * (sine / safe_c) * (1.0f - sine * sine * sine);
    const float taylor = (PI_F * 0.5f - 2.0f / 3.0f) - c;
    return c < 1.0e-6f ? taylor : regular;
}

static v3 diffuse_intensity(const closure_rec *c, v3 in, v3 out) {
    const float nl = fmaxf(vdot(c->normal, out), 0.0f);
    const v3 lambert = V3(nl * INV_PI_F, nl * INV_PI_F, nl * INV_PI_F);
    const float sigma = fclamp(c->roughness, 0.0f, 1.0f);
    if (sigma < 1.0e-5f) return lambert;
    const float a = 1.0f / (PI_F + sigma * (PI_F * 0.5f - 2.0f / 3.0f));
    const float b = sigma * a;
    const float nv = fmaxf(vdot(c->normal, in), 0.0f);
    float t = vdot(out, in) - nl * nv;
    if (t > 0.0f) t = t / (fmaxf(nl, nv) + 1.17549435e-38f);
    const float single = a + b * t;
    const v3 albedo = vclamp(c->color, 0.0f, 1.0f);
    const float two_pi = 6.2831853071795864f;
    const float e_avg = a * PI_F + ((two_pi - 5.6f) / 3.0f) * b;
    const v3 alb2 = vmul(albedo, albedo);
    const v3 e_ms = vdiv(vscale(vscale(alb2, INV_PI_F), e_avg / (1.0f - e_avg)),
                         vsub(V3(1, 1, 1), vscale(albedo, 1.0f - e_avg)));
    const float e_in = a * PI_F + b * oren_nayar_g(nv);
    const v3 ms = vscale(e_ms, 1.0f - e_in);
    const float e_out = a * PI_F + b * oren_nayar_g(nl);
    return vscale(vadd(V3(single, single, single), vscale(ms, 1.0f - e_out)), nl);
}

typedef struct eval_t {
    v3 f;
    float pdf;
    uint32_t events;
} eval_t;

enum { LOBE_DIFFUSE = EV_DIFFUSE, LOBE_TRANSPARENT = EV_TRANSPARENT, LOBE_ALL = 0x3f };

/* Mixture evaluation of every retained closure in direction `out`.
 * `selected` is the index of the sampled closure or UINT32_MAX. */
static eval_t evaluate_mixture(const closure_set *cs, v3 point_incoming, v3 ns, v3 out_dir,
                               uint32_t lobe_mask, uint32_t selected) {
    const v3 outgoing = safe_normalize(out_dir, ns);
    const v3 incoming = safe_normalize(point_incoming, vneg(outgoing));
    v3 f = V3(0, 0, 0);
    float total = 0.0f, weighted_pdf = 0.0f;
    uint32_t events = 0u;
    const int diffuse_enabled = (lobe_mask & EV_DIFFUSE) != 0u;
    const int transparent_enabled = (lobe_mask & EV_TRANSPARENT) != 0u;
    for (uint32_t i = 0; i < cs->count; ++i) {
        const closure_rec *c = &cs->c[i];
        (void)selected;
        if (c->kind == KIND_DIFFUSE) {
            /* bump shadowing is exactly one: the closure normal equals the
             * shading normal for every closure this path evaluates */
            float pdf = fmaxf(vdot(c->normal, outgoing), 0.0f) * INV_PI_F;
            const v3 value = vscale(vmul(c->weight, diffuse_intensity(c, incoming, outgoing)), 1.0f);
            const int allowed = diffuse_enabled;
            const v3 contrib = allowed ? value : V3(0, 0, 0);
            const float w = allowed ? c->sample_weight : 0.0f;
            if (allowed && sample_weight_of(contrib) > 0.0f) events |= EV_DIFFUSE | EV_REFLECTION;
            f = vadd(f, contrib);
            total += w;
            weighted_pdf += w * (allowed ? pdf : 0.0f);
        } else if (c->kind == KIND_TRANSPARENT) {
            if (transparent_enabled) total += c->sample_weight;
        }
    }
    eval_t e;
    e.f = f;
    e.pdf = total > 0.0f ? weighted_pdf / fmaxf(total, 1.0e-20f) : 0.0f;
    e.events = weighted_pdf > 0.0f ? events : 0u;
    return e;
}

typedef struct bsdf_sample_t {
    int valid;
    v3 wi;
    eval_t eval;
    uint32_t runtime_flags;
    uint32_t selected;
    float rescaled;  /* the lobe random rescaled into the chosen closure */
} bsdf_sample_t;

static uint32_t closure_runtime_flags(const closure_rec *c) {
    if (c->kind == KIND_DIFFUSE) return RT_BSDF | RT_BSDF_HAS_EVAL;
    if (c->kind == KIND_TRANSPARENT) return RT_BSDF | RT_TRANSPARENT;
    return 0u;
}

static uint32_t surface_runtime_flags(const closure_set *cs, int back_facing) {
    uint32_t f = back_facing ? RT_BACKFACING : 0u;
    for (uint32_t i = 0; i < cs->count; ++i) f |= closure_runtime_flags(&cs->c[i]);
    return f;
}

static bsdf_sample_t sample_closures(const closure_set *cs, const surface_t *sf,
                                     float u_lobe, v2 u_dir, uint32_t lobe_mask) {
    bsdf_sample_t r;
    memset(&r, 0, sizeof(r));
    r.selected = UINT32_MAX;
    const v3 incoming = safe_normalize(sf->incoming, sf->ns);
    (void)incoming;
    float weights[MAX_CLOSURES];
    float total = 0.0f;
    r.runtime_flags = sf->back_facing ? RT_BACKFACING : 0u;
    for (uint32_t i = 0; i < cs->count; ++i) {
        const closure_rec *c = &cs->c[i];
        int eligible = 0;
        if (c->kind == KIND_TRANSPARENT) eligible = (lobe_mask & EV_TRANSPARENT) != 0u;
        if (c->kind == KIND_DIFFUSE) eligible = (lobe_mask & EV_DIFFUSE) != 0u;
        weights[i] = eligible ? c->sample_weight : 0.0f;
        total += weights[i];
        r.runtime_flags |= closure_runtime_flags(c);
    }
    /* categorical inversion (surface_closure_sampling_core.cpp:410-437) */
    const float random = fclamp(u_lobe, 0.0f, 0.99999994f);
    const float target = random * total;
    float accumulated = 0.0f;
    int selected = 0;
    uint32_t sel = UINT32_MAX;
    float rescaled = 0.0f;
    for (uint32_t i = 0; i < cs->count; ++i) {
        const float next = accumulated + weights[i];
        const int choose = !selected && weights[i] > 0.0f && target < next;
        const float rs = cs->count > 1u
                             ? (target - accumulated) / fmaxf(weights[i], 1.0e-20f)
                             : random;
        if (choose) { sel = i; rescaled = rs; }
        selected |= choose;
        accumulated = next;
    }
    if (!selected) return r;
    r.selected = sel;
    r.rescaled = rescaled;
    const closure_rec *c = &cs->c[sel];
    v3 direction;
    int transparent = c->kind == KIND_TRANSPARENT;
    if (transparent) {
        direction = vneg(sf->incoming);
    } else {
        direction = sample_cosine_hemisphere(c->normal, u_dir, NULL);
    }
    const eval_t mix = evaluate_mixture(cs, sf->incoming, sf->ns, direction, lobe_mask, sel);
    const int geometric_valid = vdot(sf->ng, direction) > 0.0f;
    const int valid = transparent || geometric_valid;
    r.valid = valid;
    r.wi = direction;
    eval_t regular;
    regular.f = valid ? mix.f : V3(0, 0, 0);
    regular.pdf = valid ? mix.pdf : 0.0f;
    const v3 delta = (valid && transparent) ? vscale(c->weight, 1.0e6f) : V3(0, 0, 0);
    const float delta_num = (valid && transparent) ? 1.0e6f * weights[sel] : 0.0f;
    r.eval.f = vadd(regular.f, delta);
    r.eval.pdf = regular.pdf + delta_num / fmaxf(total, 1.0e-20f);
    uint32_t events = transparent ? (EV_TRANSMISSION | EV_TRANSPARENT)
                                  : (EV_DIFFUSE | EV_REFLECTION);
    r.eval.events = valid ? events : 0u;
    return r;
}

/* AOV (src/luisa/surface_closure_operations.cpp:387-462) */
static void surface_aov(const closure_set *cs, v3 ns, v3 *albedo, v3 *normal, v3 *transparency) {
    v3 a = V3(0, 0, 0), n = V3(0, 0, 0), tr = V3(0, 0, 0);
    float total = 0.0f;
    for (uint32_t i = 0; i < cs->count; ++i) {
        const closure_rec *c = &cs->c[i];
        if (c->kind == KIND_DIFFUSE) {
            const float pw = sample_weight_of(c->weight);
            a = vadd(a, c->weight);
            n = vadd(n, vscale(c->normal, pw));
            total += pw;
        } else if (c->kind == KIND_TRANSPARENT) {
            tr = vadd(tr, c->weight);
        }
    }
    *albedo = a;
    *transparency = tr;
    *normal = safe_normalize(total > 0.0f ? n : ns, ns);
}

static v3 transparent_extinction(const closure_set *cs) {
    v3 r = V3(0, 0, 0);
    for (uint32_t i = 0; i < cs->count; ++i)
        if (cs->c[i].kind == KIND_TRANSPARENT) r = vadd(r, cs->c[i].weight);
    return r;
}

/* ======================================================================= */
/* triangle light sampling (include/psycles/luisa/spherical_geometry.h)    */
/* ======================================================================= */

static float triangle_solid_angle(v3 ref, v3 p0, v3 p1, v3 p2) {
    const v3 a = normalize_or(vsub(p0, ref), V3(1, 0, 0));
    const v3 b = normalize_or(vsub(p1, ref), V3(0, 1, 0));
    const v3 c = normalize_or(vsub(p2, ref), V3(0, 0, 1));
    const float num = fabsf(vdot(a, vcross(b, c)));
    const float den = 1.0f + vdot(a, b) + vdot(b, c) + vdot(c, a);
    return 2.0f * atan2f(num, den);
}

static int use_solid_angle(v3 ref, v3 p0, v3 p1, v3 p2) {
    const v3 e0 = vsub(p1, p0), e1 = vsub(p2, p0), e2 = vsub(p2, p1);
    const float longest = fmaxf(vdot(e0, e0), fmaxf(vdot(e1, e1), vdot(e2, e2)));
    const v3 nn = vcross(e0, e1);
    const float nl2 = vdot(nn, nn);
    const float safe = fmaxf(nl2, 1.0e-20f);
    const float dist = vdot(nn, vsub(p0, ref)) / safe;
    return (nl2 > 1.0e-20f) && (longest > dist * dist);
}

static v2 low_distortion_barycentric(v2 r) {
    const float x = fclamp(r.x, 0.0f, 1.0f);
    const float y = fclamp(r.y, 0.0f, 1.0f);
    const int upper = y > x;
    const float u_upper = 0.5f * x;
    const float v_upper = y - u_upper;
    const float v_lower = 0.5f * y;
    const float u_lower = x - v_lower;
    v2 b = {upper ? u_upper : u_lower, upper ? v_upper : v_lower};
    return b;
}

typedef struct tri_pdf_t {
    float value;
    int valid;
} tri_pdf_t;

static tri_pdf_t triangle_directional_pdf(v3 ref, v3 lp, v3 p0, v3 p1, v3 p2) {
    const v3 e0 = vsub(p1, p0), e1 = vsub(p2, p0);
    const v3 un = vcross(e0, e1);
    const float da = sqrtf(fmaxf(vdot(un, un), 0.0f));
    const float area = 0.5f * da;
    const v3 normal = vdivs(un, fmaxf(da, 1.0e-20f));
    const v3 off = vsub(lp, ref);
    const float d2 = vdot(off, off);
    const v3 dir = vdivs(off, sqrtf(fmaxf(d2, 1.0e-20f)));
    const float cosine = fabsf(vdot(normal, vneg(dir)));
    const int solid = use_solid_angle(ref, p0, p1, p2);
    const float sa = triangle_solid_angle(ref, p0, p1, p2);
    const float area_pdf = d2 / fmaxf(cosine * area, 1.0e-20f);
    const float sa_pdf = 1.0f / fmaxf(sa, 1.0e-20f);
    const int valid_area = (area > 0.0f) && (d2 > 1.0e-12f) && (cosine > 0.0f);
    tri_pdf_t r;
    r.value = solid ? sa_pdf : area_pdf;
    r.valid = solid ? (valid_area && sa > 0.0f) : valid_area;
    return r;
}

typedef struct light_sample_t {
    v3 position, direction;
    float distance, pdf;
    int valid;
} light_sample_t;

static light_sample_t sample_triangle(v3 ref, v3 p0, v3 p1, v3 p2, v2 random) {
    const v2 ab = low_distortion_barycentric(random);
    const v3 apos = vadd(vadd(vscale(p0, 1.0f - ab.x - ab.y), vscale(p1, ab.x)), vscale(p2, ab.y));
    const v3 aoff = vsub(apos, ref);
    const float ad2 = vdot(aoff, aoff);
    const float ad = sqrtf(fmaxf(ad2, 1.0e-20f));
    const v3 adir = vdivs(aoff, ad);
    const v3 a = normalize_or(vsub(p0, ref), V3(1, 0, 0));
    const v3 b = normalize_or(vsub(p1, ref), V3(0, 1, 0));
    const v3 c = normalize_or(vsub(p2, ref), V3(0, 0, 1));
    const float cos_a = fclamp(vdot(b, c), -1.0f, 1.0f);
    const float cos_b = fclamp(vdot(a, c), -1.0f, 1.0f);
    const float cos_c = fclamp(vdot(a, b), -1.0f, 1.0f);
    (void)cos_a;
    const float sa = triangle_solid_angle(ref, p0, p1, p2);
    const v3 ab_n = normalize_or(vcross(a, b), V3(1, 0, 0));
    const v3 ac_n = normalize_or(vcross(a, c), V3(0, 1, 0));
    const float cos_alpha = fclamp(vdot(ab_n, ac_n), -1.0f, 1.0f);
    const float sin_alpha = sqrtf(fmaxf(1.0f - cos_alpha * cos_alpha, 0.0f));
    const float alpha = acosf(cos_alpha);
    const float phi = fclamp(random.x, 0.0f, 1.0f) * sa - alpha;
    const float sin_phi = sinf(phi);
    const float cos_phi = cosf(phi);
    const float u = cos_phi - cos_alpha;
    const float v = sin_phi + sin_alpha * cos_c;
    const float num = (v * cos_phi - u * sin_phi) * cos_alpha - v;
    const float den = (v * sin_phi + u * cos_phi) * sin_alpha;
    const float safe_den = fabsf(den) > 1.0e-20f ? den : (den >= 0.0f ? 1.0e-20f : -1.0e-20f);
    const float q = fabsf(den) <= 1.0e-20f ? 1.0f : num / safe_den;
    const v3 tangent = normalize_or(vsub(c, vscale(a, cos_b)), V3(0, 1, 0));
    const float qc = fclamp(q, -1.0f, 1.0f);
    const v3 c_prime = normalize_or(vadd(vscale(a, qc), vscale(tangent, sqrtf(fmaxf(1.0f - q * q, 0.0f)))), a);
    const float z = 1.0f - fclamp(random.y, 0.0f, 1.0f) * (1.0f - vdot(c_prime, b));
    const v3 edge_dir = normalize_or(vsub(c_prime, vscale(b, vdot(c_prime, b))), a);
    const v3 sdir = normalize_or(vadd(vscale(b, z), vscale(edge_dir, sqrtf(fmaxf(1.0f - z * z, 0.0f)))), b);
    const v3 e01 = vsub(p1, p0), e02 = vsub(p2, p0);
    const v3 pv = vcross(sdir, e02);
    const float det = vdot(e01, pv);
    const float safe_det = fabsf(det) > 1.0e-20f ? det : (det >= 0.0f ? 1.0e-20f : -1.0e-20f);
    const float inv_det = 1.0f / safe_det;
    const v3 tv = vsub(ref, p0);
    const float su = vdot(tv, pv) * inv_det;
    const v3 qv = vcross(tv, e01);
    const float sv = vdot(sdir, qv) * inv_det;
    const float sdist = vdot(e02, qv) * inv_det;
    const v3 spos = vadd(ref, vscale(sdir, sdist));
    const int svalid = (fabsf(det) > 1.0e-20f) && (sdist > 0.0f) && (su >= -1.0e-5f) &&
                       (sv >= -1.0e-5f) && (su + sv <= 1.0f + 1.0e-5f) && (sa > 0.0f);
    const int solid = use_solid_angle(ref, p0, p1, p2);
    light_sample_t r;
    r.position = solid ? spos : apos;
    const tri_pdf_t pdf = triangle_directional_pdf(ref, r.position, p0, p1, p2);
    r.direction = solid ? sdir : adir;
    r.distance = solid ? sdist : ad;
    r.pdf = pdf.value;
    r.valid = pdf.valid && (solid ? svalid : (ad2 > 1.0e-12f));
    return r;
}

static int samples_front(uint32_t s) {
    return s == PSX_EMISSION_SAMPLING_AUTOMATIC || s == PSX_EMISSION_SAMPLING_FRONT ||
           s == PSX_EMISSION_SAMPLING_FRONT_BACK;
}
static int samples_back(uint32_t s) {
    return s == PSX_EMISSION_SAMPLING_AUTOMATIC || s == PSX_EMISSION_SAMPLING_BACK ||
           s == PSX_EMISSION_SAMPLING_FRONT_BACK;
}
static int sampled_side(uint32_t s, v3 ng, v3 dir) {
    const int back = vdot(ng, vneg(dir)) < 0.0f;
    return back ? samples_back(s) : samples_front(s);
}

/* trace_shadow (src/luisa/path_tracer_geometry.cpp:75-174): closest hits in
 * distance order through surfaces whose transparent extinction carries
 * light, each pass restarting one float past the last hit; the first opaque
 * surface, or the transparent-bounce limit, ends the query with zero.
 * `passes` receives the traversals issued. */
/* The first surface a shadow ray evaluated (trace_shadow's first_hit,
 * src/luisa/path_tracer_geometry.cpp:136-143), for path traces. */
typedef struct shadow_first_t {
    int hit;
    uint32_t object, cprim;
    float t, u, v;
} shadow_first_t;

static v3 trace_shadow(const oracle_scene *s, ray_t sray, uint32_t visibility, uint32_t src_obj,
                       uint32_t src_prim, uint32_t light_obj, uint32_t light_prim, uint32_t tdepth,
                       uint32_t transparent_max, trace_counters *counters, uint32_t *passes,
                       shadow_first_t *first) {
    v3 transmittance = V3(1, 1, 1);
    if (first) first->hit = 0;
    for (;;) {
        const hit_t sh = closest_hit(s, &sray, visibility, src_obj, src_prim, light_obj, light_prim, counters);
        (*passes)++;
        if (!sh.valid) break;
        if (tdepth >= transparent_max) { transmittance = V3(0, 0, 0); break; }
        if (first && !first->hit) {
            first->hit = 1;
            first->object = sh.object;
            first->cprim = sh.cprim;
            first->t = sh.t;
            first->u = sh.u;
            first->v = sh.v;
        }
        const oinstance *si = &s->instances[sh.inst];
        const uint32_t smat = material_of(s, si, sh.prim);
        v3 ext = V3(0, 0, 0);
        if (smat < s->material_count) {
            surface_t ssf;
            resolve_surface(s, &sh, &sray, &ssf);
            closure_set scs;
            material_closures(&s->materials[smat], ssf.ns, &scs);
            ext = vclamp(transparent_extinction(&scs), 0.0f, 1.0f);
        }
        if (fmaxf(ext.x, fmaxf(ext.y, ext.z)) > 0.0f) {
            transmittance = vmul(transmittance, ext);
            tdepth += 1u;
            sray.tmin = bitsf(intersection_t_offset_bits(sh.t));
        } else {
            transmittance = V3(0, 0, 0);
            break;
        }
    }
    return transmittance;
}

/* ======================================================================= */
/* kernel parameters (src/luisa/path_tracer_kernel.cpp:145-374)             */
/* ======================================================================= */

typedef struct kparams {
    uint32_t window_x, window_y, window_width, window_height;
    uint32_t full_width, full_height, seed;
    uint32_t max_bounces, min_bounces, max_diffuse, max_glossy, max_transmission;
    uint32_t transparent_min, transparent_max, max_path_steps;
    uint32_t transparent_background;
    float clamp_direct, clamp_indirect;
    float light_inv_rr_threshold;
    float tan_h, tan_v, shift_x, shift_y, near_clip, far_clip;
    uint32_t projection;   /* 0 perspective, 1 orthographic */
    float ortho_span;      /* vertical view-plane span of the orthographic camera */
    float pass_alpha_threshold;
    int nee;
    int mis_mode; /* psx_direct_light_sampling actually used */
    int use_tree; /* light tree instead of the flat CDF */
    mat4 camera;
    sampler smp;
    float filter[1024];
} kparams;

static uint32_t synced(uint32_t v) { return v == UINT32_MAX ? v : v + 1u; }

static void setup_params(const oracle_scene *s, const psx_render_settings *rs,
                         const float *table, uint32_t sequence_size, kparams *k) {
    memset(k, 0, sizeof(*k));
    k->window_x = rs->window_x;
    k->window_y = rs->window_y;
    k->window_width = rs->window_width ? rs->window_width : rs->width;
    k->window_height = rs->window_height ? rs->window_height : rs->height;
    k->full_width = rs->width ? rs->width : 1u;
    k->full_height = rs->height ? rs->height : 1u;
    k->seed = rs->seed;
    const psx_integrator_settings *in = &rs->integrator;
    k->max_bounces = synced(in->max_bounces);
    k->min_bounces = synced(in->min_bounces);
    k->max_diffuse = synced(in->diffuse_bounces);
    k->max_glossy = synced(in->glossy_bounces);
    k->max_transmission = synced(in->transmission_bounces);
    k->transparent_min = synced(in->transparent_min_bounces);
    k->transparent_max = in->transparent_max_bounces;
    {
        uint64_t req = 2ull * k->max_bounces +
                       (in->transparent_max_bounces > 1u ? in->transparent_max_bounces : 1u) + 1ull;
        k->max_path_steps = (uint32_t)(req < 1024ull ? req : 1024ull);
    }
    k->transparent_background = rs->transparent_background;
    k->clamp_direct = in->sample_clamp_direct > 0.0f ? in->sample_clamp_direct * 3.0f : 0.0f;
    k->clamp_indirect = in->sample_clamp_indirect > 0.0f ? in->sample_clamp_indirect * 3.0f : 0.0f;
    k->light_inv_rr_threshold = (!in->use_light_tree && in->light_sampling_threshold > 0.0f)
                                    ? in->film_exposure / in->light_sampling_threshold
                                    : 0.0f;
    k->use_tree = in->use_light_tree != 0u;
    const int available = s->distribution_count > 0u;
    k->nee = in->direct_light_sampling != PSX_DIRECT_LIGHT_FORWARD && available;
    k->mis_mode = k->nee ? (int)in->direct_light_sampling : PSX_DIRECT_LIGHT_FORWARD;
    const float aspect = (float)k->full_width / (float)k->full_height;
    const int hfit = s->camera.sensor_fit == 0u || (s->camera.sensor_fit == 2u && aspect >= 1.0f);
    k->tan_v = hfit ? tanf(s->camera.horizontal_field_of_view * 0.5f) / aspect
                    : tanf(s->camera.field_of_view * 0.5f);
    k->tan_h = k->tan_v * aspect;
    /* orthographic_viewplane_span (include/psycles/luisa/camera_sampling.h:50-56) */
    k->projection = s->camera.projection;
    k->ortho_span = hfit ? s->camera.orthographic_scale / aspect : s->camera.orthographic_scale;
    k->shift_x = s->camera.lens_shift_x * (hfit ? 1.0f : 1.0f / aspect);
    k->shift_y = s->camera.lens_shift_y * (hfit ? aspect : 1.0f);
    k->near_clip = s->camera.near_clip;
    k->far_clip = s->camera.far_clip;
    k->pass_alpha_threshold = fclamp(rs->pass_alpha_threshold, 0.0f, 1.0f);
    memcpy(k->camera.m, s->camera.transform, sizeof(k->camera.m));
    k->smp.table = table;
    k->smp.sequence_size = sequence_size;
    oracle_pixel_filter_table(rs->pixel_filter, rs->filter_width, k->filter);
}

/* ======================================================================= */
/* light transport callables (src/luisa/path_tracer_lighting.cpp)           */
/* ======================================================================= */

static float power_heuristic(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 / fmaxf(a2 + b2, 1.0e-20f);
}

static float forward_light_weight(const kparams *k, float forward_pdf, float nee_pdf,
                                  int competing, int nee_available) {
    if (k->mis_mode == PSX_DIRECT_LIGHT_FORWARD) return 1.0f;
    if (k->mis_mode == PSX_DIRECT_LIGHT_NEE) return (competing && nee_available) ? 0.0f : 1.0f;
    return (competing && nee_available) ? power_heuristic(forward_pdf, nee_pdf) : 1.0f;
}

static float nee_light_weight(const kparams *k, float nee_pdf, float forward_pdf) {
    if (k->mis_mode == PSX_DIRECT_LIGHT_NEE) return 1.0f;
    if (k->mis_mode == PSX_DIRECT_LIGHT_FORWARD) return forward_pdf <= 0.0f ? 1.0f : 0.0f;
    return power_heuristic(nee_pdf, forward_pdf);
}

/* cycles_film_light::clamp (include/psycles/luisa/cycles_film_light.h:27-50) */
static v3 film_clamp(v3 c, uint32_t depth, float direct, float indirect) {
    c.x = isfinite(c.x) ? c.x : 0.0f;
    c.y = isfinite(c.y) ? c.y : 0.0f;
    c.z = isfinite(c.z) ? c.z : 0.0f;
    const float limit = depth > 0u ? indirect : direct;
    const float mag = fabsf(c.x) + fabsf(c.y) + fabsf(c.z);
    if (limit > 0.0f && mag > limit) return vscale(c, limit / fmaxf(mag, 1.0e-20f));
    return c;
}

/* light passes (src/luisa/path_kernel_setup.cpp:696-733,
 * src/luisa/path_tracer_lighting.cpp:107-191) */
enum { LP_DD, LP_DI, LP_GD, LP_GI, LP_TD, LP_TI, LP_EMISSION, LP_ENVIRONMENT, LP_SUMS };

static v3 component_ratio(v3 n, v3 d) {
    return V3(fabsf(d.x) > 1.0e-20f ? n.x / d.x : 0.0f, fabsf(d.y) > 1.0e-20f ? n.y / d.y : 0.0f,
              fabsf(d.z) > 1.0e-20f ? n.z / d.z : 0.0f);
}

/* split_scattered_light: diffuse, glossy and the rest (transmission) */
static void lp_scattered(v3 *lp, v3 c, v3 diffuse_w, v3 glossy_w, int direct) {
    const v3 dc = vmul(c, diffuse_w), gc = vmul(c, glossy_w);
    const v3 tc = vsub(vsub(c, dc), gc);
    lp[direct ? LP_DD : LP_DI] = vadd(lp[direct ? LP_DD : LP_DI], dc);
    lp[direct ? LP_GD : LP_GI] = vadd(lp[direct ? LP_GD : LP_GI], gc);
    lp[direct ? LP_TD : LP_TI] = vadd(lp[direct ? LP_TD : LP_TI], tc);
}

/* emission or environment seen directly, otherwise scattered light */
static void lp_route(v3 *lp, int pass, v3 c, uint32_t flag, uint32_t bounce, v3 diffuse_w, v3 glossy_w) {
    if ((flag & FLAG_ANY_PASS) == 0u)
        lp[pass] = vadd(lp[pass], c);
    else if ((flag & FLAG_SURFACE_PASS) != 0u)
        lp_scattered(lp, c, diffuse_w, glossy_w, bounce == 1u);
}

/* split_nee_light: at the first surface by the BSDF's own component
 * ratios, deeper by the path weights; every scattering closure evaluated
 * here is diffuse, so diffuse_f is the whole f and glossy_f is zero */
static void lp_nee(v3 *lp, v3 c, v3 f, uint32_t bounce, v3 path_diffuse_w, v3 path_glossy_w) {
    const int direct = bounce == 0u;
    const v3 dw = direct ? component_ratio(f, f) : path_diffuse_w;
    const v3 gw = direct ? component_ratio(V3(0, 0, 0), f) : path_glossy_w;
    lp_scattered(lp, c, dw, gw, direct);
}

static float light_roulette(v3 u, float random, float inv_threshold) {
    const float maximum = fmaxf(fabsf(u.x), fmaxf(fabsf(u.y), fabsf(u.z)));
    const float probability = maximum * inv_threshold;
    const int roulette = (inv_threshold > 0.0f) && (probability < 1.0f);
    const int survives = !roulette || (random < probability);
    const float inverse = roulette ? 1.0f / fmaxf(probability, 1.0e-20f) : 1.0f;
    return survives ? inverse : 0.0f;
}

/* ======================================================================= */
/* analytic lamps: sampling and forward intersection                       */
/* (include/psycles/luisa/analytic_light_sampling.h:84-1118,               */
/*  include/psycles/luisa/analytic_light_intersection.h:23-260,            */
/*  include/psycles/luisa/cycles_sample_mapping.h:331-387)                 */
/* ======================================================================= */
#define SM_PI 3.14159265358979323846f
#define SM_INV_PI 0.31830988618379067154f
#define SM_INV_TWO_PI 0.15915494309189533577f
#define AL_PI 3.1415926535897932f
#define AL_INV_PI 0.3183098861837907f
#define RAY_MAXIMUM 1.0e30f

/* analytic_light_sampling::safe_normalize (:120-131) */
static v3 lamp_normalize(v3 value, v3 fallback) {
    const float m2 = vdot(value, value);
    return m2 > 0.0f ? vdivs(value, sqrtf(fmaxf(m2, 1.0e-30f))) : fallback;
}

/* sin_squared_to_one_minus_cosine (cycles_sample_mapping.h:343-350) */
static float sin2_to_one_minus_cos(float s2) {
    return s2 > 0.0004f ? 1.0f - sqrtf(fmaxf(1.0f - s2, 0.0f)) : 0.5f * s2;
}

typedef struct cone_t {
    v3 direction;
    float cosine, pdf;
} cone_t;

/* sample_uniform_cone (cycles_sample_mapping.h:356-378) */
static cone_t sample_uniform_cone(v3 axis, float omc, v2 random) {
    const v2 disk = sample_uniform_disk(random);
    const float r2 = disk.x * disk.x + disk.y * disk.y;
    const float cosine = 1.0f - r2 * omc;
    const float radial = sqrtf(fmaxf(omc * (2.0f - omc * r2), 0.0f));
    const float mx = disk.x * radial, my = disk.y * radial;
    v3 t, b;
    make_orthonormals(axis, &t, &b);
    const v3 fd = vadd(vadd(vscale(t, mx), vscale(b, my)), vscale(axis, cosine));
    cone_t c;
    const int finite = omc > 0.0f;
    c.direction = finite ? fd : axis;
    c.cosine = finite ? cosine : 1.0f;
    c.pdf = finite ? SM_INV_TWO_PI / omc : 1.0f;
    return c;
}

/* sample_uniform_sphere (cycles_sample_mapping.h:380-387) */
static v3 sample_uniform_sphere(v2 r) {
    const float z = 1.0f - 2.0f * r.x;
    const float radius = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    const float phi = 2.0f * SM_PI * r.y;
    return V3(radius * cosf(phi), radius * sinf(phi), z);
}

typedef struct lamp_xform {
    v3 cx, cy, cz, ix, iy, iz;
    int valid;
} lamp_xform;

/* light_linear_transform (:137-166) */
static lamp_xform light_linear_transform(const olamp *l) {
    lamp_xform t;
    t.cx = vscale(l->axis_x, l->axis_scale.x);
    t.cy = vscale(l->axis_y, l->axis_scale.y);
    t.cz = vscale(l->axis_z, l->axis_scale.z);
    const v3 adj = vcross(t.cy, t.cz);
    const float det = vdot(t.cx, adj);
    t.valid = fabsf(det) > 1.0e-20f;
    const float inv = 1.0f / (t.valid ? det : 1.0f);
    t.ix = vscale(adj, inv);
    t.iy = vscale(vcross(t.cz, t.cx), inv);
    t.iz = vscale(vcross(t.cx, t.cy), inv);
    return t;
}

static v3 world_to_light_direction(v3 d, const lamp_xform *t) {
    return V3(vdot(t->ix, d), vdot(t->iy, d), vdot(t->iz, d));
}

typedef struct prim_hit {
    int valid;
    float distance;
    v3 position;
} prim_hit;

/* intersect_sphere (:195-238): open interval */
static prim_hit intersect_sphere(v3 ro, v3 rd, float tmin, float tmax, v3 c, float r) {
    const v3 co = vsub(c, ro);
    const float r2 = r * r;
    const float cd2 = vdot(co, co);
    const float pd = vdot(co, rd);
    const v3 perp = vsub(co, vscale(rd, pd));
    const float perp2 = vdot(perp, perp);
    const float dist = pd - copysignf(sqrtf(fmaxf(r2 - perp2, 0.0f)), cd2 - r2);
    prim_hit h;
    h.valid = !((cd2 > r2) && (pd < 0.0f)) && (perp2 <= r2) && (dist > tmin) && (dist < tmax);
    h.distance = dist;
    h.position = vadd(ro, vscale(rd, dist));
    return h;
}

/* intersect_disk (:240-273) */
static prim_hit intersect_disk(v3 ro, v3 rd, float tmin, float tmax, v3 c, v3 n, float r) {
    const v3 co = vsub(ro, c);
    const float pdist = vdot(co, n);
    const float cosine = vdot(n, vneg(rd));
    const float dist = cosine != 0.0f ? pdist / cosine : 0.0f;
    const v3 pos = vadd(ro, vscale(rd, dist));
    const v3 radial = vsub(pos, c);
    prim_hit h;
    h.valid = (pdist * cosine > 0.0f) && (dist >= 0.0f) && (vdot(radial, radial) < r * r) &&
              (dist > tmin) && (dist < tmax);
    h.distance = dist;
    h.position = pos;
    return h;
}

/* sphere_light_pdf (:281-307) */
static float sphere_light_pdf(float cd2, float r2, v3 sn, v3 dir, int had_transmission, float outside_omc) {
    const float outside = SM_INV_TWO_PI / fmaxf(outside_omc, 1.0e-30f);
    float inside;
    if (had_transmission) {
        inside = 0.5f * SM_INV_TWO_PI;
    } else {
        const float c = vdot(sn, dir);
        inside = c > 0.0f ? c * SM_INV_PI : 0.0f;
    }
    return cd2 > r2 ? outside : inside;
}

/* spot_light_to_local (:309-320) */
static v3 spot_light_to_local(v3 ray, const lamp_xform *t) {
    v3 l = lamp_normalize(world_to_light_direction(ray, t), V3(0, 0, -1));
    l.z = -l.z;
    return l;
}

/* spot_light_attenuation (:322-352) */
static float spot_light_attenuation(v3 lr, float angle, float smooth) {
    const float ch = cosf(0.5f * angle);
    const float bw = (1.0f - ch) * smooth;
    const float lin = fclamp(bw != 0.0f ? (lr.z - ch) / bw : 0.0f, 0.0f, 1.0f);
    const float sm = lin * lin * (3.0f - 2.0f * lin);
    const float hard = lr.z >= ch ? 1.0f : 0.0f;
    return bw > 0.0f ? sm : hard;
}

/* spot_one_minus_cosine_larger_spread (:354-376) */
static float spot_omc_larger_spread(float angle, v3 sc) {
    const float tn = tanf(0.5f * angle);
    const float ts = fmaxf(sc.x * sc.x, sc.y * sc.y);
    const float as = sc.z * sc.z;
    const float c = 1.0f / sqrtf(1.0f + tn * tn * ts / fmaxf(as, 1.0e-30f));
    return 1.0f - c;
}

/* point_disk_pdf (:856-873) */
static float point_disk_pdf(float d2, float light_cosine, float r) {
    const float r2 = r * r;
    const float inverse_area = r2 > 0.0f ? 1.0f / fmaxf(r2 * AL_PI, 1.0e-20f) : 1.0f;
    return light_cosine > 0.0f ? inverse_area * d2 / light_cosine : 0.0f;
}

/* point_eval_factor (:880-896) */
static float point_eval_factor(float r, int normalize) {
    const float r2 = r * r;
    const float finite_area = 4.0f * AL_PI * r2;
    const float area = r2 > 0.0f ? finite_area : 4.0f;
    const float inverse_area = normalize ? 1.0f / fmaxf(area, 1.0e-20f) : 1.0f;
    return inverse_area * AL_INV_PI;
}

typedef struct lamp_geo {
    int valid;
    v3 direction, position, normal;
    float distance, pdf;
} lamp_geo;

/* sample_sphere_geometry (:429-547) */
static lamp_geo sample_sphere_geometry(v3 ref, v3 sn, int has_transmission, v3 c, float r, v3 outside_axis,
                                       float outside_omc, int intersect_sampled, v2 rnd) {
    const float r2 = r * r;
    v3 ln = vsub(ref, c);
    const float cd2 = vdot(ln, ln);
    const float cd = sqrtf(fmaxf(cd2, 1.0e-30f));
    ln = vdivs(ln, cd);
    const int outside = cd2 > r2;
    v3 dir = V3(0, 0, 0);
    float cosine = 1.0f, pdf = 0.0f, dist = 0.0f;
    v3 pos = ref;
    int known = 0;
    int valid = (r > 0.0f) && (cd2 > 0.0f);
    if (outside) {
        const cone_t cone = sample_uniform_cone(outside_axis, outside_omc, rnd);
        dir = cone.direction;
        cosine = cone.cosine;
        pdf = cone.pdf;
        if (intersect_sampled) {
            const prim_hit h = intersect_sphere(ref, dir, 0.0f, RAY_MAXIMUM, c, r);
            valid = valid && h.valid;
            dist = h.distance;
            pos = h.position;
            known = h.valid;
        }
    } else {
        if (has_transmission) {
            dir = sample_uniform_sphere(rnd);
            pdf = 0.5f * SM_INV_TWO_PI;
        } else {
            float hp;
            dir = sample_cosine_hemisphere(sn, rnd, &hp);
            pdf = hp;
        }
        cosine = -vdot(dir, ln);
    }
    if (!known) {
        dist = cd * cosine - copysignf(sqrtf(fmaxf(r2 - cd2 + cd2 * cosine * cosine, 0.0f)), cd2 - r2);
        pos = vadd(ref, vscale(dir, dist));
    }
    lamp_geo g;
    g.normal = lamp_normalize(vsub(pos, c), vneg(dir));
    g.position = vadd(vscale(g.normal, r), c);
    g.valid = valid;
    g.direction = dir;
    g.distance = dist;
    g.pdf = pdf;
    return g;
}

/* sample_disk_geometry (:549-603) */
static lamp_geo sample_disk_geometry(v3 ref, v3 c, float r, v2 rnd) {
    v3 ln = vsub(ref, c);
    const float cd2 = vdot(ln, ln);
    const float cd = sqrtf(fmaxf(cd2, 1.0e-30f));
    ln = vdivs(ln, cd);
    v3 t, b;
    make_orthonormals(ln, &t, &b);
    const v2 disk = sample_uniform_disk(rnd);
    const v3 pos = vadd(c, vscale(vadd(vscale(t, disk.x), vscale(b, disk.y)), r));
    const v3 off = vsub(pos, ref);
    const float d2 = vdot(off, off);
    const float d = sqrtf(fmaxf(d2, 1.0e-30f));
    const v3 dir = vdivs(off, d);
    const float cosine = vdot(ln, vneg(dir));
    lamp_geo g;
    g.valid = (cd2 > 0.0f) && (cosine > 0.0f);
    g.direction = dir;
    g.position = pos;
    g.normal = vneg(dir);
    g.distance = d;
    g.pdf = point_disk_pdf(d2, cosine, r);
    return g;
}

typedef struct lamp_sample_t {
    int valid;
    v3 direction, position, normal;
    float distance, pdf, eval;
} lamp_sample_t;

/* sample_point_light (:898-979) and sample_spot_light (:981-1105) */
/* segment: the spot proposal of a volume segment
 * (VolumeAnalyticLightSampling::spot_from_segment,
 * src/luisa/volume_analytic_light_sampling.cpp): the visible sphere cap
 * outside a finite sphere, never the spread cone, and no reject of a zero
 * attenuation */
static lamp_sample_t sample_point_or_spot_ex(const olamp *l, v3 ref, v3 sn, int has_transmission, v2 rnd,
                                             int segment) {
    const int spot = l->type == PSX_LIGHT_SPOT;
    const int sphere = (l->flags & LAMP_SPHERE) != 0u;
    const int normalize = (l->flags & LAMP_NORMALIZE) != 0u;
    const v3 co = vsub(ref, l->position);
    const float cd2 = vdot(co, co);
    const float r2 = l->radius * l->radius;
    const v3 light_normal = lamp_normalize(co, V3(0, 0, -1));
    const float sphere_cap = sin2_to_one_minus_cos(r2 / fmaxf(cd2, 1.0e-30f));
    v3 outside_axis = vneg(light_normal);
    float outside_cap = sphere_cap;
    int spread_cone = 0;
    if (spot && !segment) {
        const float spot_cap = spot_omc_larger_spread(l->spot_angle, l->axis_scale);
        spread_cone = (cd2 > r2) && !(sphere_cap < spot_cap);
        if (spread_cone) {
            outside_axis = l->axis_z;
            outside_cap = spot_cap;
        }
    }
    const int finite_sphere = sphere && (l->radius > 0.0f);
    lamp_geo g;
    if (finite_sphere)
        g = sample_sphere_geometry(ref, sn, has_transmission, l->position, l->radius, outside_axis, outside_cap,
                                   spread_cone, rnd);
    else
        g = sample_disk_geometry(ref, l->position, l->radius, rnd);
    lamp_sample_t s;
    s.direction = g.direction;
    s.position = g.position;
    s.normal = g.normal;
    s.distance = g.distance;
    s.pdf = g.pdf;
    s.eval = point_eval_factor(l->radius, normalize);
    s.valid = g.valid;
    if (spot) {
        const lamp_xform xf = light_linear_transform(l);
        const v3 local = spot_light_to_local(vneg(g.direction), &xf);
        const float attenuation = spot_light_attenuation(local, l->spot_angle, l->spot_smooth);
        const int use_attenuation = !finite_sphere || (cd2 > r2);
        s.eval = s.eval * (use_attenuation ? attenuation : 1.0f);
        if (!segment) s.valid = g.valid && (s.eval != 0.0f);
    }
    return s;
}

static lamp_sample_t sample_point_or_spot(const olamp *l, v3 ref, v3 sn, int has_transmission, v2 rnd) {
    return sample_point_or_spot_ex(l, ref, sn, has_transmission, rnd, 0);
}

/* sample_distant_light (:84-108) */
static lamp_sample_t sample_distant_light(const olamp *l, v2 rnd) {
    const float half_angle = 0.5f * fmaxf(l->angle, 0.0f);
    const float sq = sinf(0.5f * half_angle);
    const float omc = 2.0f * sq * sq;
    const cone_t ns = sample_uniform_cone(vneg(l->axis_z), omc, rnd);
    const int finite = half_angle > 0.0f;
    const float sh = sinf(half_angle);
    const float disk_area = AL_PI * sh * sh;
    lamp_sample_t s;
    s.direction = vneg(ns.direction);
    s.pdf = ns.pdf;
    s.eval = ((l->flags & LAMP_NORMALIZE) != 0u && finite) ? 1.0f / fmaxf(disk_area, 1.0e-20f) : 1.0f;
    s.position = vneg(s.direction);
    s.normal = vneg(s.direction);
    s.distance = RAY_MAXIMUM;
    s.valid = 1;
    return s;
}

typedef struct lamp_hit_t {
    int valid;
    float distance, pdf, eval;
} lamp_hit_t;

/* intersect_point / intersect_spot (analytic_light_intersection.h:33-260) */
static lamp_hit_t intersect_lamp(const olamp *l, v3 ro, v3 rd, float tmin, float tmax, v3 previous_normal,
                                 int had_transmission) {
    const int spot = l->type == PSX_LIGHT_SPOT;
    const int sphere = (l->flags & LAMP_SPHERE) != 0u;
    const v3 co = vsub(ro, l->position);
    const float cd2 = vdot(co, co);
    const float r2 = l->radius * l->radius;
    const float sphere_cap = sin2_to_one_minus_cos(r2 / fmaxf(cd2, 1.0e-30f));
    float outside_cap = sphere_cap;
    if (spot) outside_cap = fminf(sphere_cap, spot_omc_larger_spread(l->spot_angle, l->axis_scale));
    /* intersect_point_geometry (:33-117) */
    const v3 light_normal = lamp_normalize(co, vneg(rd));
    prim_hit g;
    float pdf;
    v3 normal = vneg(rd);
    if (sphere) {
        g = intersect_sphere(ro, rd, tmin, tmax, l->position, l->radius);
        normal = lamp_normalize(vsub(g.position, l->position), vneg(rd));
        pdf = sphere_light_pdf(cd2, r2, previous_normal, rd, had_transmission, outside_cap);
    } else {
        g = intersect_disk(ro, rd, tmin, tmax, l->position, light_normal, l->radius);
        const float light_cosine = vdot(light_normal, vneg(rd));
        pdf = point_disk_pdf(g.distance * g.distance, light_cosine, l->radius);
    }
    (void)normal;
    lamp_hit_t h;
    h.valid = g.valid && (l->radius > 0.0f);
    h.distance = g.distance;
    h.pdf = pdf;
    h.eval = point_eval_factor(l->radius, (l->flags & LAMP_NORMALIZE) != 0u);
    if (spot) {
        const lamp_xform xf = light_linear_transform(l);
        const v3 local = spot_light_to_local(vneg(rd), &xf);
        const int use_attenuation = !sphere || (cd2 > r2);
        const float attenuation = spot_light_attenuation(local, l->spot_angle, l->spot_smooth);
        h.eval = h.eval * (use_attenuation ? attenuation : 1.0f);
        h.valid = h.valid && (vdot(rd, co) < 0.0f) && (h.eval > 0.0f);
    } else {
        h.valid = h.valid && (h.eval > 0.0f);
    }
    return h;
}

/* ---- area lamps (src/luisa/area_light_sampling.cpp,
 * include/psycles/luisa/analytic_light_sampling.h:605-846) ---------------- */
static float al_sdiv(float a, float b) { return b != 0.0f ? a / b : 0.0f; }

/* area_spread_attenuation (:796-846) */
static float area_spread_attenuation(v3 dir_to_light, v3 light_normal, float spread) {
    const float cosine = fmaxf(vdot(light_normal, vneg(dir_to_light)), 0.0f);
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

/* rectangle_solid_angle_context (:605-690): the rectangle seen from `ref`
 * in the frame of its edges (Urena, Fajardo, King, "An area-preserving
 * parametrization for spherical rectangles", EGSR 2013) */
typedef struct rect_ctx {
    v3 z;
    float z0, x0, x1, y0, y1;
    float e[4], a[4];
    float solid_angle, pdf;
} rect_ctx;

static rect_ctx rect_solid_angle_context(v3 ref, v3 center, v3 au, float lu, v3 av, float lv) {
    rect_ctx c;
    v3 z = vcross(au, av);
    const v3 off = vsub(center, ref);
    float z0 = vdot(off, z);
    if (z0 > 0.0f) {
        z = vneg(z);
        z0 = -z0;
    }
    c.z = z;
    c.z0 = z0;
    const float cu = vdot(off, au), cv = vdot(off, av);
    c.x0 = cu - 0.5f * lu;
    c.x1 = cu + 0.5f * lu;
    c.y0 = cv - 0.5f * lv;
    c.y1 = cv + 0.5f * lv;
    const float raw[4] = {-c.y0, c.x1, c.y1, -c.x0};
    for (int i = 0; i < 4; ++i) c.e[i] = raw[i] / sqrtf(raw[i] * raw[i] + z0 * z0);
    for (int i = 0; i < 4; ++i) {
        const float p = -c.e[i] * c.e[(i + 1) & 3];
        c.a[i] = asinf(fminf(fmaxf(p, -1.0f), 1.0f));
    }
    c.solid_angle = -(c.a[0] + c.a[1] + c.a[2] + c.a[3]);
    const float min_e2 = fminf(fminf(c.e[0] * c.e[0], c.e[1] * c.e[1]), fminf(c.e[2] * c.e[2], c.e[3] * c.e[3]));
    const float dist = vlength(off);
    const float planar_pdf = al_sdiv(-(dist * dist * dist), z0 * lu * lv);
    const int planar = (c.solid_angle < 1.0e-5f) || (min_e2 > 0.99999f);
    c.pdf = planar ? planar_pdf : 1.0f / c.solid_angle;
    return c;
}

/* sample_rectangle_solid_angle (:716-794) */
static v3 sample_rect_solid_angle(v3 ref, v3 center, v3 au, float lu, v3 av, float lv, v2 rnd, float *pdf) {
    const rect_ctx c = rect_solid_angle_context(ref, center, au, lu, av, lv);
    const float b0 = c.e[0], b1 = c.e[2];
    const float b0sq = b0 * b0;
    const float alpha = rnd.x * c.solid_angle + c.a[2] + c.a[3];
    const float fu = al_sdiv(cosf(alpha) * b0 + b1, sinf(alpha));
    float cos_u = copysignf(1.0f / sqrtf(fu * fu + b0sq), fu);
    cos_u = fminf(fmaxf(cos_u, -1.0f), 1.0f);
    float su = -(cos_u * c.z0) / fmaxf(sqrtf(fmaxf(1.0f - cos_u * cos_u, 0.0f)), 1.0e-7f);
    su = fminf(fmaxf(su, c.x0), c.x1);
    const float du2 = su * su + c.z0 * c.z0;
    const float h0 = c.y0 / sqrtf(du2 + c.y0 * c.y0);
    const float h1 = c.y1 / sqrtf(du2 + c.y1 * c.y1);
    const float hv = h0 + rnd.y * (h1 - h0);
    const float hv2 = hv * hv;
    const float sv = hv2 < 1.0f - 1.0e-6f ? hv * sqrtf(du2 / (1.0f - hv2)) : c.y1;
    *pdf = c.pdf;
    return vadd(vadd(vadd(ref, vscale(au, su)), vscale(av, sv)), vscale(c.z, c.z0));
}

static float area_value(const olamp *l) {
    float area = l->size_u * l->size_v;
    if ((l->flags & LAMP_ELLIPSE) != 0u) area *= 0.25f * AL_PI;
    return area;
}

static float area_inverse_area(const olamp *l) {
    const float area = area_value(l);
    return ((l->flags & LAMP_NORMALIZE) != 0u && area != 0.0f) ? al_sdiv(1.0f, area) : 1.0f;
}

static float area_eval_factor(const olamp *l, v3 direction, v3 normal) {
    return AL_INV_PI * area_inverse_area(l) * area_spread_attenuation(direction, normal, l->spread);
}

static float area_tangent_half_spread(float spread) {
    return spread == AL_PI ? FLT_MAX : tanf(0.5f * fmaxf(spread, 0.0f));
}

static float area_to_solid_angle(v3 normal, v3 direction, float distance_squared) {
    const float cosine = vdot(normal, vneg(direction));
    return cosine > 0.0f ? distance_squared / cosine : 0.0f;
}

/* clamp_to_spread (area_light_sampling.cpp:179-420): the part of the lamp
 * that the spread cone about the reference point's foot can see, as a
 * rectangle, a circle or the original shape */
typedef struct area_shape {
    int valid, rectangle;
    v3 center, au, av;
    float lu, lv;
} area_shape;

static area_shape area_original_shape(const olamp *l) {
    area_shape r;
    r.valid = 1;
    r.rectangle = (l->flags & LAMP_ELLIPSE) == 0u;
    r.center = l->position;
    r.au = l->axis_x;
    r.av = l->axis_y;
    r.lu = l->size_u;
    r.lv = l->size_v;
    return r;
}

static area_shape area_clamp_to_spread(const olamp *l, v3 ref, v3 normal, float tangent) {
    area_shape r = area_original_shape(l);
    const float plane_distance = vdot(normal, vsub(ref, l->position));
    const v3 closest = vsub(ref, vscale(normal, plane_distance));
    const float spread_radius = plane_distance * tangent;
    const float spread_u = vdot(l->axis_x, vsub(closest, l->position));
    const float spread_v = vdot(l->axis_y, vsub(closest, l->position));
    const int round = !r.rectangle && r.lu == r.lv;
    int sample_spread = spread_radius == 0.0f;
    int resolved = 0;
    if (round && !sample_spread) {
        const float center_distance = sqrtf(spread_u * spread_u + spread_v * spread_v);
        const float light_radius = 0.5f * r.lu;
        if (center_distance >= light_radius + spread_radius) {
            r.valid = 0;
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
                    r.rectangle = 1;
                    r.au = safe_normalize(vsub(l->position, closest), l->axis_x);
                    r.av = vcross(normal, r.au);
                    r.lu = rect_u;
                    r.lv = rect_v;
                    r.center = vscale(vadd(vadd(l->position, closest), vscale(r.au, spread_radius - light_radius)),
                                      0.5f);
                    resolved = 1;
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
            r.valid = 0;
        } else {
            const float rect_u = max_u - min_u, rect_v = max_v - min_v;
            const float rect_area = rect_u * rect_v;
            const float ellipse_area = r.rectangle ? FLT_MAX : 0.25f * AL_PI * r.lu * r.lv;
            const float spread_area = AL_PI * spread_radius * spread_radius;
            if (r.rectangle || rect_area < fminf(ellipse_area, spread_area)) {
                const float cu = 0.5f * (min_u + max_u), cv = 0.5f * (min_v + max_v);
                r.center = vadd(r.center, vadd(vscale(r.au, cu), vscale(r.av, cv)));
                r.lu = rect_u;
                r.lv = rect_v;
                r.rectangle = 1;
                resolved = 1;
            } else {
                r.rectangle = 0;
                sample_spread = spread_area < ellipse_area;
            }
        }
    }
    if (r.valid && !resolved && sample_spread) {
        r.rectangle = 0;
        r.center = vadd(l->position, vadd(vscale(r.au, spread_u), vscale(r.av, spread_v)));
        r.lu = 2.0f * spread_radius;
        r.lv = 2.0f * spread_radius;
    }
    return r;
}

static v3 area_sample_ellipse(v3 center, v3 au, float lu, v3 av, float lv, v2 rnd) {
    const v2 d = sample_uniform_disk(rnd);
    return vadd(center, vadd(vscale(au, 0.5f * lu * d.x), vscale(av, 0.5f * lv * d.y)));
}

/* AreaLightSampling::from_position (:494-646): surface NEE */
static lamp_sample_t sample_area_light(const olamp *l, v3 ref, v2 rnd) {
    lamp_sample_t s;
    memset(&s, 0, sizeof(s));
    const v3 normal = vneg(l->axis_z);
    s.position = l->position;
    s.normal = normal;
    if (!(vdot(vsub(l->position, ref), normal) <= 0.0f)) return s;
    const float tangent = area_tangent_half_spread(l->spread);
    area_shape sh = area_original_shape(l);
    if ((l->flags & LAMP_FULL_SPREAD) == 0u) sh = area_clamp_to_spread(l, ref, normal, tangent);
    if (!sh.valid) return s;
    v3 position = sh.center;
    float pdf = 0.0f;
    if (sh.rectangle) {
        position = sample_rect_solid_angle(ref, sh.center, sh.au, sh.lu, sh.av, sh.lv, rnd, &pdf);
    } else if (tangent == 0.0f) {
        pdf = 1.0f;
    } else {
        position = area_sample_ellipse(sh.center, sh.au, sh.lu, sh.av, sh.lv, rnd);
        pdf = 4.0f * AL_INV_PI / (sh.lu * sh.lv);
    }
    const v3 off = vsub(position, ref);
    const float d2 = vdot(off, off);
    const float d = sqrtf(fmaxf(d2, 0.0f));
    const v3 dir = vdivs(off, fmaxf(d, 1.0e-30f));
    if (!sh.rectangle && tangent > 0.0f) pdf *= area_to_solid_angle(normal, dir, d2);
    const float factor = area_eval_factor(l, dir, normal);
    const v3 inplane = vsub(position, l->position);
    const float lu = vdot(inplane, l->axis_x), lv = vdot(inplane, l->axis_y);
    const float eu = (0.5f + 1.0e-7f) * l->size_u + 1.0e-6f;
    const float ev = (0.5f + 1.0e-7f) * l->size_v + 1.0e-6f;
    const int inside_rect = fabsf(lu) <= eu && fabsf(lv) <= ev;
    const int inside_ellipse = lu * lu / (eu * eu) + lv * lv / (ev * ev) <= 1.0f;
    s.valid = ((l->flags & LAMP_ELLIPSE) ? inside_ellipse : inside_rect) && factor > 0.0f;
    s.direction = dir;
    s.position = position;
    s.distance = d;
    s.pdf = pdf;
    s.eval = factor;
    return s;
}

/* intersect_area (analytic_light_intersection.h:271-353) with
 * AreaLightSampling::from_intersection (area_light_sampling.cpp:648-726) */
static lamp_hit_t intersect_area_lamp(const olamp *l, v3 ro, v3 rd, float tmin, float tmax) {
    lamp_hit_t h;
    const v3 normal = vneg(l->axis_z);
    const float dn = vdot(rd, normal);
    const float distance = al_sdiv(vdot(vsub(l->position, ro), normal), dn);
    const v3 position = vadd(ro, vscale(rd, distance));
    const v3 inplane = vsub(position, l->position);
    const float u = al_sdiv(vdot(inplane, l->axis_x), l->size_u);
    const float v = al_sdiv(vdot(inplane, l->axis_y), l->size_v);
    const int inside_rect = u >= -0.5f && u <= 0.5f && v >= -0.5f && v <= 0.5f;
    const int inside = inside_rect && (!(l->flags & LAMP_ELLIPSE) || u * u + v * v <= 0.25f);
    const float tangent = area_tangent_half_spread(l->spread);
    area_shape sh = area_original_shape(l);
    if ((l->flags & LAMP_FULL_SPREAD) == 0u) sh = area_clamp_to_spread(l, ro, normal, tangent);
    float pdf = 0.0f, factor = 0.0f;
    int eval_valid = 0;
    if (sh.valid) {
        if (sh.rectangle) {
            pdf = rect_solid_angle_context(ro, sh.center, sh.au, sh.lu, sh.av, sh.lv).pdf;
        } else if (tangent == 0.0f) {
            pdf = 1.0f;
        } else {
            pdf = 4.0f * AL_INV_PI / (sh.lu * sh.lv);
            pdf *= area_to_solid_angle(normal, rd, distance * distance);
        }
        factor = area_eval_factor(l, rd, normal);
        eval_valid = area_value(l) > 0.0f && factor > 0.0f;
    }
    h.valid = dn < 0.0f && distance > tmin && distance < tmax && inside && eval_valid;
    h.distance = distance;
    h.pdf = pdf;
    h.eval = factor;
    return h;
}

int oracle_sample_lamp(const psx_light *light, const float reference[3], const float shading_normal[3],
                       int has_transmission, const float random[2], float out[13]) {
    if (!light || light->type > PSX_LIGHT_DISTANT) return -1;
    olamp l;
    lamp_from_desc(light, &l);
    const v2 rnd = {random[0], random[1]};
    const lamp_sample_t ls = l.type == PSX_LIGHT_DISTANT ? sample_distant_light(&l, rnd)
                             : l.type == PSX_LIGHT_AREA  ? sample_area_light(&l, load3(reference), rnd)
                                                         : sample_point_or_spot(&l, load3(reference),
                                                                                load3(shading_normal),
                                                                                has_transmission, rnd);
    const float v[13] = {ls.direction.x, ls.direction.y, ls.direction.z, ls.pdf, ls.position.x, ls.position.y,
                         ls.position.z, ls.distance, ls.normal.x, ls.normal.y, ls.normal.z, ls.eval,
                         ls.valid ? 1.0f : 0.0f};
    memcpy(out, v, sizeof(v));
    return 0;
}

int oracle_intersect_lamp(const psx_light *light, const float origin[3], const float direction[3], float tmin,
                          float tmax, const float previous_normal[3], int had_transmission, float out[4]) {
    if (!light || (light->type != PSX_LIGHT_POINT && light->type != PSX_LIGHT_SPOT && light->type != PSX_LIGHT_AREA))
        return -1;
    olamp l;
    lamp_from_desc(light, &l);
    const lamp_hit_t h = l.type == PSX_LIGHT_AREA
                             ? intersect_area_lamp(&l, load3(origin), load3(direction), tmin, tmax)
                             : intersect_lamp(&l, load3(origin), load3(direction), tmin, tmax,
                                              load3(previous_normal), had_transmission);
    out[0] = h.valid ? 1.0f : 0.0f;
    out[1] = h.distance;
    out[2] = h.pdf;
    out[3] = h.eval;
    return 0;
}

/* make_triangle_emitter_callable (src/luisa/path_tracer_light_tree.cpp:
 * 746-790): emitter id of an emissive (object, Cycles primitive) */
static uint32_t lt_triangle_emitter(const oracle_scene *s, uint32_t object, uint32_t cprim) {
    uint32_t first = 0u, length = s->emitter_count;
    while (length > 0u) {
        const uint32_t half = length >> 1u;
        const uint32_t middle = first + half;
        const uint32_t eo = s->emitters[middle].instance, ep = s->emitter_cprim[middle];
        if (eo < object || (eo == object && ep < cprim)) {
            first = middle + 1u;
            length = length - half - 1u;
        } else {
            length = half;
        }
    }
    if (first < s->emitter_count && s->emitters[first].instance == object && s->emitter_cprim[first] == cprim)
        return first;
    return LT_INVALID;
}

int oracle_light_tree_query(const oracle_scene *s, const float point[3], const float normal[3],
                            int has_transmission, float u, uint32_t *emitter_id, float *pdf_sample,
                            float *pdf_reverse) {
    if (!s) return -1;
    const v3 p = load3(point), n = load3(normal);
    float pdf = 0.0f;
    const uint32_t id = lt_sample(&s->tree, u, p, n, has_transmission, &pdf);
    *emitter_id = id;
    *pdf_sample = pdf;
    *pdf_reverse = id == LT_INVALID ? 0.0f : lt_pdf(&s->tree, id, p, n, has_transmission);
    return 0;
}

float oracle_light_tree_pdf(const oracle_scene *s, uint32_t emitter_id, const float point[3],
                            const float normal[3], int has_transmission) {
    return s ? lt_pdf(&s->tree, emitter_id, load3(point), load3(normal), has_transmission) : 0.0f;
}

uint32_t oracle_light_tree_emitters(const oracle_scene *s) { return s ? s->tree.emitter_count : 0u; }

int oracle_light_tree_sample(const oracle_light_tree *tree, const float point[3], const float normal[3],
                             int has_transmission, float u, uint32_t *emitter_id, float *pdf_sample,
                             float *pdf_reverse) {
    if (!tree) return -1;
    const v3 p = load3(point), n = load3(normal);
    float pdf = 0.0f;
    const uint32_t id = lt_sample(tree, u, p, n, has_transmission, &pdf);
    *emitter_id = id;
    *pdf_sample = pdf;
    *pdf_reverse = id == LT_INVALID ? 0.0f : lt_pdf(tree, id, p, n, has_transmission);
    return 0;
}

float oracle_light_tree_reverse_pdf(const oracle_light_tree *tree, uint32_t emitter_id, const float point[3],
                                    const float normal[3], int has_transmission) {
    return tree ? lt_pdf(tree, emitter_id, load3(point), load3(normal), has_transmission) : 0.0f;
}

/* ======================================================================= */
/* the path integrator (src/luisa/path_kernel_pipeline.cpp:73-185 and the  */
/* stages it calls)                                                        */
/* ======================================================================= */

typedef struct film_pixel {
    float combined[4], normal[4], albedo[4];
    uint32_t count;
} film_pixel;

typedef struct trace_log {
    uint32_t *records; /* PSX_TRACE_EVENTS * PSX_TRACE_WORDS, or NULL */
    uint32_t events;
    float *slots;      /* PSX_PATH_TRACE_SLOTS * 4 schema v2 slots, or NULL */
} trace_log;

/* Diagnostic path trace record, schema "psycles.cycles-path-trace" v2
 * (include/psycles/luisa/path_trace_schema.h:10-121): 8 global slots, then
 * 72 slots per surface event for the first 4 events; w = 1 marks a written
 * slot.  trace_uint32 (src/luisa/path_kernel_setup.cpp:430-433) stores a
 * word as its two 16-bit halves. */
#define TRACE_GLOBAL_SLOTS 8u
#define TRACE_EVENT_SLOTS 72u
#define TRACE_MAX_EVENTS 4u

static void trace_slot(float *slots, uint32_t index, float x, float y, float z) {
    float *d = slots + 4u * index;
    d[0] = x;
    d[1] = y;
    d[2] = z;
    d[3] = 1.0f;
}
static void trace_slot3(float *slots, uint32_t index, v3 v) { trace_slot(slots, index, v.x, v.y, v.z); }
static float lo16(uint32_t v) { return (float)(v & 0xffffu); }
static float hi16(uint32_t v) { return (float)(v >> 16); }
#define TRACE_SHADOW_BASE 296u
#define TRACE_SHADOW_SLOTS 8u

/* slots of surface event `step`, or NULL past the traced events */
static float *event_slots(trace_log *log, uint32_t step) {
    return log && log->slots && step < TRACE_MAX_EVENTS
               ? log->slots + 4u * (TRACE_GLOBAL_SLOTS + step * TRACE_EVENT_SLOTS)
               : NULL;
}
/* shadow-event slots of event `step`, or NULL */
static float *shadow_event_slots(trace_log *log, uint32_t step) {
    return log && log->slots && step < TRACE_MAX_EVENTS
               ? log->slots + 4u * (TRACE_SHADOW_BASE + step * TRACE_SHADOW_SLOTS)
               : NULL;
}
/* an object or primitive identity, -1 when invalid
 * (src/luisa/path_kernel_direct_light_trace.cpp:6-11) */
static float trace_identity(uint32_t v) { return v != UINT32_MAX ? (float)v : -1.0f; }
/* Cycles' exclusion flags of an object's ray visibility, not a shadow
 * catcher (src/luisa/cycles_shader_identity.h:64-97) */
static uint32_t object_visibility_flags(uint32_t v) {
    return ((v & PSX_VISIBILITY_CAMERA) ? 0u : (1u << 24)) | ((v & PSX_VISIBILITY_DIFFUSE) ? 0u : (1u << 27)) |
           ((v & PSX_VISIBILITY_GLOSSY) ? 0u : (1u << 26)) |
           ((v & PSX_VISIBILITY_TRANSMISSION) ? 0u : (1u << 25)) |
           ((v & PSX_VISIBILITY_VOLUME_SCATTER) ? 0u : (1u << 23)) | (1u << 22);
}
/* Cycles KernelLightType (src/luisa/cycles_shader_identity.h:37-53) */
static uint32_t cycles_light_type(uint32_t type) {
    return type == PSX_LIGHT_POINT ? 0u : type == PSX_LIGHT_DISTANT ? 1u : type == PSX_LIGHT_AREA ? 3u : 4u;
}
/* Cycles ClosureType (src/luisa/graph_surface_scattering.cpp:373-420,
 * include/psycles/luisa/cycles_closure.h:18-32) */
static uint32_t cycles_closure_type(const closure_rec *c) {
    if (c->kind == KIND_TRANSPARENT) return 30u;
    return c->roughness < 1.0e-5f ? 2u : 3u;
}
/* record_shadow and record_contribution
 * (src/luisa/path_kernel_direct_light_trace.cpp:152-228): the traced shadow
 * ray, its identities, first hit and transmittance; the unclamped
 * contribution when light arrives.  `unshadowed` is throughput x surviving. */
static void trace_shadow_record(trace_log *log, uint32_t step, const ray_t *r, int cast_shadow, uint32_t src_obj,
                                uint32_t src_prim, int skip_self, uint32_t light_obj, uint32_t light_prim,
                                const shadow_first_t *first, v3 transmittance, v3 unshadowed) {
    float *e = shadow_event_slots(log, step);
    if (!e) return;
    trace_slot3(e, 0u, r->o);
    trace_slot3(e, 1u, r->d);
    trace_slot(e, 2u, r->tmin, r->tmax, cast_shadow ? 1.0f : 0.0f);
    trace_slot(e, 3u, trace_identity(src_obj), trace_identity(src_prim), skip_self ? 1.0f : 0.0f);
    trace_slot(e, 4u, trace_identity(light_obj), trace_identity(light_prim), 0.0f);
    if (first->hit) {
        trace_slot(e, 5u, trace_identity(first->object), trace_identity(first->cprim), 0.0f);
        trace_slot(e, 6u, first->t, first->u, first->v);
    } else {
        trace_slot(e, 5u, -1.0f, -1.0f, -1.0f);
        trace_slot(e, 6u, 0.0f, 0.0f, 0.0f);
    }
    trace_slot3(e, 7u, transmittance);
    if (transmittance.x > 0.0f || transmittance.y > 0.0f || transmittance.z > 0.0f)
        trace_slot3(event_slots(log, step), 47u, vmul(unshadowed, transmittance));
}

static void log_event(trace_log *log, uint32_t step, uint32_t word, uint32_t value) {
    if (!log || !log->records || step >= PSX_TRACE_EVENTS) return;
    log->records[step * PSX_TRACE_WORDS + word] = value;
}
static void log_flag(trace_log *log, uint32_t step, uint32_t bit) {
    if (!log || !log->records || step >= PSX_TRACE_EVENTS) return;
    log->records[step * PSX_TRACE_WORDS + PSX_TRACE_FLAGS] |= bit;
}

typedef struct camera_ray_t {
    ray_t ray;
    uint32_t rng_hash;
} camera_ray_t;

/* sample_camera_dimensions + construct_camera_ray (perspective and
 * orthographic; src/luisa/path_tracer_camera.cpp:10-244) */
static camera_ray_t camera_ray(const kparams *k, uint32_t full_x, uint32_t full_y,
                               uint32_t sample_index) {
    camera_ray_t r;
    const uint32_t cycles_y = k->full_height - 1u - full_y;
    r.rng_hash = oracle_pixel_hash(full_x, cycles_y, k->seed);
    const float *f = sample_4d(&k->smp, sample_index, r.rng_hash, 0u);
    v2 filter = {f[0], f[1]};
    if (sample_index == 0u) filter.x = filter.y = 0.5f;
    const float jx = oracle_pixel_filter_sample(k->filter, filter.x);
    const float jy = 1.0f - oracle_pixel_filter_sample(k->filter, filter.y);
    const float width = (float)k->full_width;
    const float height = (float)k->full_height;
    const float sx = 2.0f * ((float)full_x + jx) / width - 1.0f + 2.0f * k->shift_x;
    const float sy = 1.0f - 2.0f * ((float)full_y + jy) / height + 2.0f * k->shift_y;
    v3 local_o = V3(0, 0, 0), local_dir = V3(0, 0, -1);
    float clip_cos = 1.0f;
    if (k->projection == 1u) {
        /* orthographic (src/luisa/path_tracer_camera.cpp:127-143) */
        local_o = V3(sx * k->ortho_span * (width / height) * 0.5f, sy * k->ortho_span * 0.5f, 0.0f);
    } else {
        local_dir = vnormalize(V3(sx * k->tan_h, sy * k->tan_v, -1.0f));
        clip_cos = -local_dir.z;
    }
    const v3 world_o = xform_point(&k->camera, local_o);
    const v3 world_d = xform_dir(&k->camera, local_dir);
    const v3 d = safe_normalize(world_d, V3(0, 0, -1));
    float clip[2];
    oracle_camera_clip_range(k->near_clip, k->far_clip, clip_cos, clip);
    r.ray.o = vadd(world_o, vscale(d, clip[0]));
    r.ray.d = d;
    r.ray.tmin = 0.0f;
    r.ray.tmax = clip[1];
    return r;
}

typedef struct path_out {
    v3 radiance, normal, albedo;
    float transparency;
    uint64_t rays;
    v3 lp[LP_SUMS];
} path_out;

static v3 triangle_world_position_mm(const oinstance *inst, v3 p) {
    return mat_mul(&inst->o2w, p, 1.0f);
}

static void integrate(const oracle_scene *s, const kparams *k, uint32_t full_x,
                      uint32_t full_y, uint32_t sample_index, path_out *out,
                      trace_log *log, trace_counters *counters) {
    const camera_ray_t cam = camera_ray(k, full_x, full_y, sample_index);
    const uint32_t rng_hash = cam.rng_hash;
    ray_t ray = cam.ray;
    if (log && log->slots) {
        /* global slots (src/luisa/path_kernel_setup.cpp:540-560) */
        const float *f = sample_4d(&k->smp, sample_index, rng_hash, 0u);
        const float fx = sample_index == 0u ? 0.5f : f[0], fy = sample_index == 0u ? 0.5f : f[1];
        trace_slot(log->slots, 0u, 2.0f, (float)full_x, (float)(k->full_height - 1u - full_y));
        trace_slot(log->slots, 1u, (float)sample_index, lo16(rng_hash), hi16(rng_hash));
        trace_slot(log->slots, 2u, fx, fy, 0.0f);
        trace_slot(log->slots, 3u, 0.0f, 0.0f, 0.0f);
        trace_slot3(log->slots, 4u, ray.o);
        trace_slot3(log->slots, 5u, ray.d);
        trace_slot(log->slots, 6u, ray.tmin, ray.tmax, 0.5f);
    }
    uint32_t src_obj = UINT32_MAX, src_prim = UINT32_MAX;
    uint32_t ray_visibility = PSX_VISIBILITY_CAMERA;
    v3 radiance = V3(0, 0, 0), throughput = V3(1, 1, 1);
    v3 sample_normal = V3(0, 0, 0), sample_albedo = V3(0, 0, 0);
    v3 lp[LP_SUMS];
    for (int i = 0; i < LP_SUMS; ++i) lp[i] = V3(0, 0, 0);
    v3 path_diffuse_w = V3(0, 0, 0), path_glossy_w = V3(0, 0, 0);
    float transparency = 0.0f;
    int primary_recorded = 0;
    float previous_bsdf_pdf = 0.0f;
    float minimum_bsdf_pdf = FLT_MAX;
    int previous_delta = 1;
    v3 previous_mis_origin_normal = V3(0, 0, 0);
    uint32_t ray_events = 0u;
    oracle_path_state ps;
    oracle_initial_state(&ps);
    int terminate_after_transparent = 0, terminate_on_next_surface = 0;
    uint64_t rays = 0;
    const oracle_path_limits lim = {k->max_bounces, k->max_diffuse, k->max_glossy,
                                    k->max_transmission, k->transparent_max};

    for (uint32_t step = 0; step < k->max_path_steps; ++step) {
        log_flag(log, step, 1u);
        log_event(log, step, PSX_TRACE_RNG_OFFSET, ps.rng_offset);
        log_event(log, step, PSX_TRACE_LIGHT, UINT32_MAX);
        log_event(log, step, PSX_TRACE_CLOSURE, UINT32_MAX);
        log_event(log, step, PSX_TRACE_OBJECT, UINT32_MAX);
        log_event(log, step, PSX_TRACE_PRIMITIVE, UINT32_MAX);
        /* bounce setup (src/luisa/path_kernel_bounce_setup.cpp:19-152) */
        const float u_term = sample_4d(&k->smp, sample_index, rng_hash, ps.rng_offset + 0u)[0];
        const float *ul = sample_4d(&k->smp, sample_index, rng_hash, ps.rng_offset + 1u);
        const v3 u_light = V3(ul[0], ul[1], ul[2]);
        const dist_entry *light = NULL;
        uint32_t light_id = UINT32_MAX;
        dist_entry tree_light;
        if (s->distribution_count > 0u && !k->use_tree) {
            light = distribution_sample(s->distribution, s->distribution_count, u_light.z);
            light_id = (uint32_t)(light - s->distribution);
        }
        const float u_lightterm = sample_4d(&k->smp, sample_index, rng_hash, ps.rng_offset + 2u)[0];
        const float *ub = sample_4d(&k->smp, sample_index, rng_hash, ps.rng_offset + 3u);
        const v3 u_bsdf = V3(ub[0], ub[1], ub[2]);
        log_event(log, step, PSX_TRACE_TERMINATE_BITS, fbits(u_term));
        const hit_t hit = closest_hit(s, &ray, ray_visibility, src_obj, src_prim, UINT32_MAX,
                                      UINT32_MAX, counters);
        rays++;
        /* Lamps are transparent closest events resolved before the mesh or
         * background event, without Sobol dimensions
         * (src/luisa/path_kernel_pipeline.cpp:79-137,
         *  src/luisa/path_kernel_closest_event.cpp:24-266,
         *  src/luisa/path_kernel_forward_light.cpp:59-117) */
        int lamp_terminated = 0;
        {
            uint32_t previous_lamp = UINT32_MAX;
            const float surface_distance = hit.valid ? hit.t : ray.tmax;
            const int camera_path = (ray_visibility & PSX_VISIBILITY_CAMERA) != 0u;
            const int had_transmission = (ps.flag & FLAG_MIS_HAD_TRANSMISSION) != 0u;
            for (;;) {
                float event_distance = surface_distance;
                uint32_t found = UINT32_MAX;
                lamp_hit_t best;
                for (uint32_t li = 0; li < s->lamp_count; ++li) {
                    const olamp *l = &s->lamps[li];
                    if (l->type == PSX_LIGHT_DISTANT) continue;
                    const int eligible = (l->visibility & ray_visibility) != 0u &&
                                         (camera_path || (l->flags & LAMP_USE_MIS) != 0u) && li != previous_lamp;
                    if (!eligible) continue;
                    const lamp_hit_t c = l->type == PSX_LIGHT_AREA
                                             ? intersect_area_lamp(l, ray.o, ray.d, ray.tmin, event_distance)
                                             : intersect_lamp(l, ray.o, ray.d, ray.tmin, event_distance,
                                                              previous_mis_origin_normal, had_transmission);
                    if (c.valid) {
                        found = li;
                        best = c;
                        event_distance = c.distance;
                    }
                }
                if (found == UINT32_MAX) break;
                const olamp *l = &s->lamps[found];
                const float selection_pdf =
                    k->use_tree ? lt_pdf(&s->tree, s->emitter_count + found, ray.o, previous_mis_origin_normal,
                                         had_transmission)
                                : s->light_selection_pdf;
                const float light_pdf = best.pdf * selection_pdf;
                const int competing = ps.bounce > 0u && !previous_delta;
                const float mis = forward_light_weight(k, previous_bsdf_pdf, light_pdf, competing, light_pdf > 0.0f);
                const v3 light_radiance = vscale(l->color, l->power * best.eval);
                const v3 contribution = film_clamp(vscale(vmul(throughput, light_radiance), mis),
                                                   ps.bounce > 1u ? 1u : 0u, k->clamp_direct, k->clamp_indirect);
                radiance = vadd(radiance, contribution);
                lp_route(lp, LP_EMISSION, contribution, ps.flag, ps.bounce, path_diffuse_w, path_glossy_w);
                ps.transparent_bounce += 1u;
                ray.tmin = bitsf(intersection_t_offset_bits(best.distance));
                previous_lamp = found;
                if (ps.transparent_bounce >= k->transparent_max) {
                    lamp_terminated = 1;
                    break;
                }
            }
        }
        if (lamp_terminated) break;
        if (!hit.valid) {
            /* background event (src/luisa/path_kernel_background_event.cpp:20-159) */
            log_flag(log, step, 4u);
            const v3 contribution = film_clamp(vscale(vmul(throughput, s->background), 1.0f),
                                               ps.bounce > 1u ? 1u : 0u, k->clamp_direct,
                                               k->clamp_indirect);
            const int world_visible = (s->world_visibility & ray_visibility) != 0u;
            const int transparent_bg = k->transparent_background &&
                                       (ps.flag & FLAG_TRANSPARENT_BACKGROUND) != 0u;
            if (transparent_bg) {
                transparency += (throughput.x + throughput.y + throughput.z) * (1.0f / 3.0f);
            } else if (world_visible) {
                radiance = vadd(radiance, contribution);
            }
            lp_route(lp, LP_ENVIRONMENT, world_visible ? contribution : V3(0, 0, 0), ps.flag, ps.bounce,
                     path_diffuse_w, path_glossy_w);
            break;
        }
        log_flag(log, step, 2u);
        log_event(log, step, PSX_TRACE_OBJECT, hit.object);
        log_event(log, step, PSX_TRACE_PRIMITIVE, hit.cprim);
        /* surface geometry */
        surface_t sf;
        resolve_surface(s, &hit, &ray, &sf);
        const omaterial *mat = sf.material < s->material_count ? &s->materials[sf.material] : NULL;
        closure_set cs;
        if (mat) material_closures(mat, sf.ns, &cs);
        else { cs.count = 0; cs.emission = V3(0, 0, 0); }
        uint32_t lobe_mask = LOBE_ALL;
        if (terminate_after_transparent) lobe_mask = EV_TRANSPARENT;
        /* surface shading (src/luisa/path_kernel_surface_shading.cpp:115-250) */
        const v3 emitted = cs.emission;
        float emission_weight = 1.0f;
        if (k->nee && s->emitter_count > 0u &&
            sf.emission_sampling != PSX_EMISSION_SAMPLING_NONE) {
            const int competing = ps.bounce > 0u && !previous_delta;
            const v3 oriented = sf.back_facing ? vneg(sf.ng) : sf.ng;
            float selection_pdf =
                0.5f * vlength(vcross(vsub(sf.wp1, sf.wp0), vsub(sf.wp2, sf.wp0))) *
                s->triangle_area_pdf;
            if (k->use_tree)
                selection_pdf = lt_pdf(&s->tree, lt_triangle_emitter(s, sf.object, sf.cprim), ray.o,
                                       previous_mis_origin_normal, (ps.flag & FLAG_MIS_HAD_TRANSMISSION) != 0u);
            const tri_pdf_t tp = triangle_directional_pdf(ray.o, sf.position, sf.wp0, sf.wp1, sf.wp2);
            const v3 off = vsub(sf.position, ray.o);
            const v3 dir = normalize_or(off, V3(0, 0, 1));
            const int side = sampled_side(sf.emission_sampling, oriented, dir);
            const float area = 0.5f * vlength(vcross(vsub(sf.wp1, sf.wp0), vsub(sf.wp2, sf.wp0)));
            const float value = tp.value * selection_pdf;
            const int valid = side && tp.valid && area > 0.0f && value > 0.0f;
            const float light_pdf = valid ? value : 0.0f;
            emission_weight = forward_light_weight(k, previous_bsdf_pdf, light_pdf, competing,
                                                   light_pdf > 0.0f);
        }
        const v3 emission_contribution =
            film_clamp(vscale(vmul(throughput, emitted), emission_weight),
                       ps.bounce > 1u ? 1u : 0u, k->clamp_direct, k->clamp_indirect);
        radiance = vadd(radiance, emission_contribution);
        lp_route(lp, LP_EMISSION, emission_contribution, ps.flag, ps.bounce, path_diffuse_w, path_glossy_w);
        if (terminate_on_next_surface) break;
        /* Russian roulette */
        float continuation = 1.0f;
        {
            const float th[3] = {throughput.x, throughput.y, throughput.z};
            const float p = oracle_continuation_probability(ps.flag, ps.bounce, ps.transparent_bounce,
                                                            k->min_bounces, k->transparent_min, th);
            continuation = p;
            log_event(log, step, PSX_TRACE_RR_BITS, fbits(p));
            if (p <= 0.0f) { log_flag(log, step, 8u); break; }
            if (p != 1.0f) {
                if (u_term >= p) { log_flag(log, step, 8u); break; }
                throughput = vdivs(throughput, p);
            }
            /* surface event slots once the path survived the roulette
             * (src/luisa/path_kernel_surface_shading.cpp:270-340) */
            if (log && log->slots && step < TRACE_MAX_EVENTS) {
                float *e = log->slots + 4u * (TRACE_GLOBAL_SLOTS + step * TRACE_EVENT_SLOTS);
                trace_slot(e, 0u, (float)step, (float)ps.rng_offset, (float)ps.bounce);
                trace_slot(e, 1u, (float)ps.transparent_bounce, (float)ps.diffuse_bounce, (float)ps.glossy_bounce);
                trace_slot(e, 2u, (float)ps.transmission_bounce, lo16(ps.visibility), hi16(ps.visibility));
                trace_slot(e, 3u, lo16(ps.flag), hi16(ps.flag), p);
                trace_slot3(e, 4u, throughput);
                trace_slot3(e, 5u, ray.o);
                trace_slot3(e, 6u, ray.d);
                trace_slot(e, 7u, ray.tmin, ray.tmax, 0.5f);
                trace_slot(e, 8u, hit.t, hit.u, hit.v);
                trace_slot(e, 9u, (float)hit.object, (float)hit.cprim, 1.0f);
                trace_slot(e, 10u, lo16(sf.material), hi16(sf.material), (float)cs.count);
                trace_slot3(e, 11u, sf.position);
                trace_slot3(e, 12u, sf.ng);
                trace_slot3(e, 13u, sf.ns);
                const uint32_t rf = surface_runtime_flags(&cs, sf.back_facing);
                trace_slot(e, 38u, lo16(rf), hi16(rf), 0.0f);
                trace_slot(e, 14u, u_term, k->light_inv_rr_threshold > 0.0f ? u_lightterm : 0.0f, 0.0f);
                trace_slot3(e, 15u, u_light);
                trace_slot3(e, 16u, u_bsdf);
                /* the shading point's closures (src/luisa/path_kernel_surface_shading.cpp:350-380) */
                for (uint32_t c = 0; c < cs.count && c < 8u; ++c) {
                    trace_slot(e, 48u + 3u * c, (float)c, (float)cycles_closure_type(&cs.c[c]), cs.c[c].sample_weight);
                    trace_slot3(e, 49u + 3u * c, cs.c[c].weight);
                    trace_slot3(e, 50u + 3u * c, cs.c[c].normal);
                }
            }
        }
        if (ps.bounce == 0u) {
            v3 aov_albedo, aov_normal, aov_tr;
            surface_aov(&cs, sf.ns, &aov_albedo, &aov_normal, &aov_tr);
            sample_albedo = vadd(sample_albedo, vmul(throughput, aov_albedo));
            const v3 alpha = vclamp(vsub(V3(1, 1, 1), aov_tr), 0.0f, 1.0f);
            const float avg_alpha = (alpha.x + alpha.y + alpha.z) * (1.0f / 3.0f);
            const int writes = !primary_recorded &&
                               (k->pass_alpha_threshold == 0.0f || avg_alpha >= k->pass_alpha_threshold);
            if (writes) sample_normal = aov_normal;
            primary_recorded |= writes;
            if (writes) ps.flag |= FLAG_SINGLE_PASS_DONE;
        }
        const uint32_t runtime_flags = surface_runtime_flags(&cs, sf.back_facing);
        /* light tree selection (src/luisa/path_kernel_pipeline.cpp:140-150) */
        if (k->use_tree) {
            float tree_pdf = 0.0f;
            const uint32_t id = lt_sample(&s->tree, u_light.z, sf.position, sf.ns,
                                          (runtime_flags & RT_BSDF_HAS_TRANSMISSION) != 0u, &tree_pdf);
            if (id != LT_INVALID && id < s->distribution_count) {
                tree_light = s->distribution[id];
                tree_light.selection_pdf = tree_pdf;
                light = &tree_light;
                light_id = id;
            }
        }
        /* next-event estimation: emissive mesh component
         * (src/luisa/path_kernel_nee_emissive_mesh.cpp:30-298) */
        if (k->nee && (runtime_flags & RT_BSDF_HAS_EVAL) != 0u && light && light->kind == 0u) {
            log_flag(log, step, 16u);
            log_event(log, step, PSX_TRACE_LIGHT, light_id);
            const emitter *em = &s->emitters[light->index];
            const oinstance *einst = &s->instances[em->instance];
            const omesh *emesh = &s->meshes[einst->mesh];
            v3 lp0, lp1, lp2;
            tri_positions(emesh->positions, emesh->tris, em->prim, &lp0, &lp1, &lp2);
            const v3 p0 = triangle_world_position_mm(einst, lp0);
            const v3 p1 = triangle_world_position_mm(einst, lp1);
            const v3 p2 = triangle_world_position_mm(einst, lp2);
            const v3 local_n = normalize_or(vcross(vsub(lp1, lp0), vsub(lp2, lp0)), V3(0, 0, 1));
            const v3 egn = normalize_or(mat_mul(&einst->n2w, local_n, 0.0f), V3(0, 0, 1));
            const float earea = 0.5f * vlength(vcross(vsub(p1, p0), vsub(p2, p0)));
            const v2 rnd = {u_light.x, u_light.y};
            const light_sample_t ls = sample_triangle(sf.position, p0, p1, p2, rnd);
            const int side = sampled_side(em->emission_sampling, egn, ls.direction);
            const float lpdf = ls.pdf * light->selection_pdf;
            const int lvalid = ls.valid && side && earea > 0.0f && lpdf > 0.0f;
            float *te = event_slots(log, step);
            if (te) {
                /* record_sample / record_failed_sample
                 * (src/luisa/path_kernel_direct_light_trace.cpp:55-105) */
                if (lvalid) {
                    const int esmooth = emesh->smooth && emesh->smooth[em->prim] != 0u;
                    const uint32_t shader = em->material | (1u << 30) | (esmooth ? (1u << 31) : 0u) | (1u << 28) |
                                            object_visibility_flags(einst->visibility & 0x3fu);
                    trace_slot(te, 17u, 5.0f, (float)light_id, (float)(emesh->prim_offset + em->prim));
                    trace_slot(te, 18u, (float)em->instance, 0.0f, 0.0f);
                    trace_slot(te, 40u, lo16(shader), hi16(shader), 0.0f);
                    trace_slot(te, 19u, lpdf, light->selection_pdf, 1.0f);
                    trace_slot3(te, 20u, ls.direction);
                    trace_slot3(te, 21u, ls.position);
                    trace_slot3(te, 22u, egn);
                } else {
                    trace_slot(te, 17u, -1.0f, -1.0f, -1.0f);
                }
            }
            if (lvalid) {
                const uint32_t eobj = em->instance;
                const uint32_t eprim = emesh->prim_offset + em->prim;
                const int is_transmission = vdot(ls.direction, sf.ns) < 0.0f;
                const int same = eprim == sf.cprim && eobj == sf.object;
                const v3 oriented = is_transmission ? vneg(sf.ng) : sf.ng;
                const int reject_self = same && vdot(ls.direction, oriented) > 0.0f;
                if (!reject_self) {
                    const v3 le = constant_emission(&s->materials[em->material]);
                    const eval_t ev = evaluate_mixture(&cs, sf.incoming, sf.ns, ls.direction,
                                                       lobe_mask, UINT32_MAX);
                    const float mis = nee_light_weight(k, lpdf, ev.pdf);
                    const v3 weighted = vscale(vmul(ev.f, le), mis / lpdf);
                    if (te) {
                        trace_slot(te, 23u, ls.distance, ev.pdf, mis);
                        trace_slot3(te, 41u, ev.f);
                        trace_slot3(te, 42u, ev.f);
                        trace_slot(te, 43u, 0.0f, 0.0f, 0.0f);
                        trace_slot3(te, 44u, weighted);
                    }
                    if (vany_nonzero(weighted)) {
                        const v3 unshadowed = weighted;
                        const float rr = light_roulette(unshadowed, u_lightterm, k->light_inv_rr_threshold);
                        const v3 surviving = vscale(unshadowed, rr);
                        if (te) {
                            trace_slot(te, 45u, 1.0f, 1.0f, 1.0f);
                            trace_slot3(te, 46u, vmul(throughput, surviving));
                        }
                        if (vany_nonzero(surviving)) {
                            const shadow_origin_t sso = surface_shadow_origin(s, &sf, ls.direction);
                            const v3 so = sso.position;
                            const uint32_t sh_obj = sso.skip_self ? sf.object : UINT32_MAX;
                            const uint32_t sh_prim = sso.skip_self ? sf.cprim : UINT32_MAX;
                            const v3 soff = vsub(ls.position, so);
                            const float sdist = sqrtf(fmaxf(vdot(soff, soff), 1.0e-20f));
                            ray_t sray;
                            sray.o = so;
                            sray.d = vdivs(soff, sdist);
                            sray.tmin = 0.0f;
                            sray.tmax = sdist;
                            /* trace_shadow (src/luisa/path_tracer_geometry.cpp:75-174) */
                            uint32_t passes = 0u;
                            shadow_first_t first;
                            const v3 transmittance = trace_shadow(s, sray, PSX_VISIBILITY_SHADOW, sh_obj, sh_prim,
                                                                  eobj, eprim, ps.transparent_bounce,
                                                                  k->transparent_max, counters, &passes, &first);
                            rays += passes;
                            trace_shadow_record(log, step, &sray, 1, sh_obj, sh_prim, sso.skip_self, eobj, eprim,
                                                &first, transmittance, vmul(throughput, surviving));
                            if (transmittance.x > 0.0f || transmittance.y > 0.0f || transmittance.z > 0.0f) {
                                const v3 c = film_clamp(vmul(vmul(throughput, surviving), transmittance),
                                                        ps.bounce, k->clamp_direct, k->clamp_indirect);
                                radiance = vadd(radiance, c);
                                lp_nee(lp, c, ev.f, ps.bounce, path_diffuse_w, path_glossy_w);
                            }
                        }
                    }
                }
            }
        }
        /* next-event estimation: analytic lamps
         * (src/luisa/path_kernel_nee_analytic.cpp:98-410) */
        if (k->nee && (runtime_flags & RT_BSDF_HAS_EVAL) != 0u && light && light->kind == 1u) {
            log_flag(log, step, 16u);
            log_event(log, step, PSX_TRACE_LIGHT, light_id);
            const olamp *l = &s->lamps[light->index];
            const int reached_max_bounces = ps.bounce > l->max_bounces;
            const v2 rnd = {u_light.x, u_light.y};
            const int distant = l->type == PSX_LIGHT_DISTANT;
            const int has_transmission = (runtime_flags & RT_BSDF_HAS_TRANSMISSION) != 0u;
            const lamp_sample_t ls = distant ? sample_distant_light(l, rnd)
                                     : l->type == PSX_LIGHT_AREA
                                         ? sample_area_light(l, sf.position, rnd)
                                         : sample_point_or_spot(l, sf.position, sf.ns, has_transmission, rnd);
            const v3 light_radiance = vscale(l->color, l->power * ls.eval);
            const float light_pdf = ls.pdf * light->selection_pdf;
            float *te = event_slots(log, step);
            if (te) {
                /* record_sample / record_failed_sample
                 * (src/luisa/path_kernel_nee_analytic.cpp:233-249, 413) */
                if (!reached_max_bounces && ls.valid && light_pdf > 0.0f) {
                    trace_slot(te, 17u, (float)cycles_light_type(l->type), (float)light_id, (float)light->index);
                    trace_slot(te, 18u, (float)UINT32_MAX, 0.0f, 0.0f);
                    trace_slot(te, 40u, lo16(UINT32_MAX), hi16(UINT32_MAX), 0.0f);
                    trace_slot(te, 19u, light_pdf, light->selection_pdf, ls.eval);
                    trace_slot3(te, 20u, ls.direction);
                    trace_slot3(te, 21u, ls.position);
                    trace_slot3(te, 22u, ls.normal);
                } else {
                    trace_slot(te, 17u, -1.0f, -1.0f, -1.0f);
                }
            }
            if (!reached_max_bounces && ls.valid && light_pdf > 0.0f) {
                const eval_t ev = evaluate_mixture(&cs, sf.incoming, sf.ns, ls.direction, lobe_mask, UINT32_MAX);
                const float cycles_mis = nee_light_weight(k, light_pdf, ev.pdf);
                const float mis = (l->flags & LAMP_FORWARD) != 0u ? cycles_mis : 1.0f;
                const v3 weighted = vscale(vmul(ev.f, light_radiance), mis / fmaxf(light_pdf, 1.0e-20f));
                if (te) {
                    trace_slot(te, 23u, ls.distance, ev.pdf, cycles_mis);
                    trace_slot3(te, 41u, ev.f);
                    trace_slot3(te, 42u, ev.f);
                    trace_slot(te, 43u, 0.0f, 0.0f, 0.0f);
                    trace_slot3(te, 44u, weighted);
                }
                if (vany_nonzero(weighted)) {
                    const float rr = light_roulette(weighted, u_lightterm, k->light_inv_rr_threshold);
                    const v3 surviving = vscale(weighted, rr);
                    if (te) {
                        trace_slot(te, 45u, 1.0f, 1.0f, 1.0f);
                        trace_slot3(te, 46u, vmul(throughput, surviving));
                    }
                    if (vany_nonzero(surviving)) {
                        const shadow_origin_t sso = surface_shadow_origin(s, &sf, ls.direction);
                        const v3 so = sso.position;
                        const uint32_t sh_obj = sso.skip_self ? sf.object : UINT32_MAX;
                        const uint32_t sh_prim = sso.skip_self ? sf.cprim : UINT32_MAX;
                        const v3 foff = vsub(ls.position, so);
                        const float fdist = sqrtf(fmaxf(vdot(foff, foff), 1.0e-20f));
                        ray_t sray;
                        sray.o = so;
                        sray.d = distant ? ls.direction : vdivs(foff, fdist);
                        sray.tmin = 0.0f;
                        sray.tmax = distant ? RAY_MAXIMUM : fdist;
                        uint32_t passes = 0u;
                        shadow_first_t first;
                        const v3 transmittance = trace_shadow(s, sray, PSX_VISIBILITY_SHADOW, sh_obj, sh_prim,
                                                              UINT32_MAX, UINT32_MAX, ps.transparent_bounce,
                                                              k->transparent_max, counters, &passes, &first);
                        rays += passes;
                        trace_shadow_record(log, step, &sray, (l->flags & LAMP_CAST_SHADOW) != 0u, sh_obj, sh_prim,
                                            sso.skip_self, UINT32_MAX, UINT32_MAX, &first, transmittance,
                                            vmul(throughput, surviving));
                        if (transmittance.x > 0.0f || transmittance.y > 0.0f || transmittance.z > 0.0f) {
                            const v3 c = film_clamp(vmul(vmul(throughput, surviving), transmittance), ps.bounce,
                                                    k->clamp_direct, k->clamp_indirect);
                            radiance = vadd(radiance, c);
                            lp_nee(lp, c, ev.f, ps.bounce, path_diffuse_w, path_glossy_w);
                        }
                    }
                }
            }
        }
        /* surface scatter (src/luisa/path_kernel_surface_scatter.cpp:16-324) */
        const v2 udir = {u_bsdf.x, u_bsdf.y};
        const bsdf_sample_t bs = sample_closures(&cs, &sf, u_bsdf.z, udir, lobe_mask);
        log_event(log, step, PSX_TRACE_CLOSURE, bs.selected);
        {
            /* the sampled closure and BSDF sample
             * (src/luisa/path_kernel_surface_scatter.cpp:95-130) */
            float *te = event_slots(log, step);
            if (te && bs.selected < cs.count) {
                const closure_rec *c = &cs.c[bs.selected];
                trace_slot(te, 24u, (float)bs.selected, (float)cycles_closure_type(c), c->sample_weight);
                trace_slot(te, 25u, u_bsdf.x, u_bsdf.y, bs.rescaled);
                trace_slot3(te, 26u, c->weight);
                trace_slot3(te, 27u, c->normal);
                trace_slot(te, 28u, bs.eval.pdf, bs.eval.pdf, (float)label_from_events(bs.eval.events));
                trace_slot3(te, 29u, bs.wi);
                trace_slot3(te, 30u, bs.eval.f);
                const float r = bs.valid && c->kind == KIND_DIFFUSE ? 1.0f : 0.0f;
                trace_slot(te, 31u, r, r, 1.0f);
            }
        }
        if (!bs.valid || bs.eval.pdf <= 0.0f) break;
        log_flag(log, step, 32u);
        const int transparent = (bs.eval.events & EV_TRANSPARENT) != 0u;
        const int singular = (bs.eval.events & EV_SINGULAR) != 0u;
        if (transparent) log_flag(log, step, 64u);
        const uint32_t label = label_from_events(bs.eval.events);
        if (ps.bounce == 0u && !transparent) {
            /* src/luisa/path_kernel_surface_scatter.cpp:163-176 */
            path_diffuse_w = component_ratio(bs.eval.f, bs.eval.f);
            path_glossy_w = component_ratio(V3(0, 0, 0), bs.eval.f);
        }
        throughput = vmul(throughput, vdivs(bs.eval.f, bs.eval.pdf));
        if (vany_nan(throughput) || throughput.x < 0.0f || throughput.y < 0.0f || throughput.z < 0.0f)
            break;
        ray_events = transparent ? (ray_events | bs.eval.events) : bs.eval.events;
        oracle_path_state next;
        oracle_next_surface(&ps, label, bs.runtime_flags, &lim, &next);
        ps = next;
        ray_visibility = oracle_contract_visibility(ps.visibility);
        terminate_on_next_surface = (ps.flag & FLAG_TERMINATE_ON_NEXT_SURFACE) != 0u;
        terminate_after_transparent = (ps.flag & FLAG_TERMINATE_AFTER_TRANSPARENT) != 0u;
        if (!transparent) {
            previous_bsdf_pdf = bs.eval.pdf;
            minimum_bsdf_pdf = fminf(minimum_bsdf_pdf, bs.eval.pdf);
            previous_delta = singular;
            previous_mis_origin_normal = sf.ns;
        }
        const v3 nd = vnormalize(bs.wi);
        if (transparent) {
            ray.tmin = bitsf(intersection_t_offset_bits(hit.t));
        } else {
            ray.o = surface_ray_origin(s, &sf, nd);
            ray.d = nd;
            ray.tmin = 0.0f;
            ray.tmax = 1.0e30f;
        }
        src_obj = sf.object;
        src_prim = sf.cprim;
        /* post-scatter slots (src/luisa/path_kernel_surface_scatter.cpp:293-320) */
        if (log && log->slots && step < TRACE_MAX_EVENTS) {
            float *e = log->slots + 4u * (TRACE_GLOBAL_SLOTS + step * TRACE_EVENT_SLOTS);
            trace_slot(e, 32u, (float)ps.bounce, (float)ps.transparent_bounce, (float)ps.rng_offset);
            trace_slot3(e, 33u, throughput);
            trace_slot3(e, 34u, ray.o);
            trace_slot3(e, 35u, ray.d);
            trace_slot(e, 36u, lo16(ps.flag), hi16(ps.flag), 0.0f);
            trace_slot(e, 37u, previous_bsdf_pdf, minimum_bsdf_pdf, continuation);
            trace_slot(e, 39u, lo16(ps.visibility), hi16(ps.visibility), 0.0f);
        }
    }
    (void)minimum_bsdf_pdf;
    (void)ray_events;
    out->radiance = radiance;
    out->normal = sample_normal;
    out->albedo = sample_albedo;
    out->transparency = transparency;
    out->rays = rays;
    for (int i = 0; i < LP_SUMS; ++i) out->lp[i] = lp[i];
}

/* ======================================================================= */
/* rendering, threading and passes                                          */
/* ======================================================================= */

typedef struct job {
    const oracle_scene *scene;
    const kparams *k;
    uint32_t first, count;
    uint32_t row_begin, row_end;
    float *combined, *normal, *albedo;
    uint32_t *sample_count;
    float *light; /* ORACLE_LIGHT_PASSES float4 per pixel, or NULL */
    uint32_t next_row;
    pthread_mutex_t lock;
    uint64_t rays, nodes, tris;
} job;

static void render_pixel(job *j, uint32_t px, uint32_t py, trace_counters *tc, uint64_t *rays) {
    const kparams *k = j->k;
    const size_t index = (size_t)py * k->window_width + px;
    float *c = j->combined + 4 * index;
    float *n = j->normal + 4 * index;
    float *a = j->albedo + 4 * index;
    uint32_t completed = j->sample_count[index];
    for (uint32_t si = 0; si < j->count; ++si) {
        path_out o;
        integrate(j->scene, k, px + k->window_x, py + k->window_y, j->first + si, &o, NULL, tc);
        *rays += o.rays;
        v3 r = o.radiance;
        if (vany_nan(r)) r = V3(0, 0, 0);
        c[0] += r.x; c[1] += r.y; c[2] += r.z; c[3] += o.transparency;
        n[0] += o.normal.x; n[1] += o.normal.y; n[2] += o.normal.z; n[3] += 1.0f;
        a[0] += o.albedo.x; a[1] += o.albedo.y; a[2] += o.albedo.z; a[3] += 1.0f;
        if (j->light) {
            /* LightPassBuffer order (src/luisa/path_tracer_internal.h:410-424) */
            const v3 zero = V3(0, 0, 0);
            const v3 buf[ORACLE_LIGHT_PASSES] = {o.lp[LP_DD], o.lp[LP_DI], o.lp[LP_GD], o.lp[LP_GI],
                                                 o.lp[LP_TD], o.lp[LP_TI], zero, zero,
                                                 o.lp[LP_EMISSION], o.lp[LP_ENVIRONMENT], zero, zero};
            float *l = j->light + 4u * ORACLE_LIGHT_PASSES * index;
            for (int b = 0; b < ORACLE_LIGHT_PASSES; ++b) {
                l[4 * b] += buf[b].x;
                l[4 * b + 1] += buf[b].y;
                l[4 * b + 2] += buf[b].z;
                l[4 * b + 3] += 1.0f;
            }
        }
        completed += 1u;
    }
    j->sample_count[index] = completed;
}

static void *worker(void *arg) {
    job *j = (job *)arg;
    trace_counters tc = {0, 0};
    uint64_t rays = 0;
    for (;;) {
        pthread_mutex_lock(&j->lock);
        const uint32_t row = j->next_row++;
        pthread_mutex_unlock(&j->lock);
        if (row >= j->row_end) break;
        for (uint32_t x = 0; x < j->k->window_width; ++x) render_pixel(j, x, row, &tc, &rays);
    }
    pthread_mutex_lock(&j->lock);
    j->rays += rays;
    j->nodes += tc.nodes;
    j->tris += tc.tris;
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

int oracle_render_rows(const oracle_scene *s, const psx_render_settings *rs, uint32_t first,
                       uint32_t count, uint32_t total, int threads, uint32_t row_begin,
                       uint32_t row_end, float *combined, float *normal, float *albedo,
                       uint32_t *sample_count, oracle_stats *stats) {
    return oracle_render_light(s, rs, first, count, total, threads, row_begin, row_end, combined, normal, albedo,
                               sample_count, NULL, stats);
}

int oracle_render_light(const oracle_scene *s, const psx_render_settings *rs, uint32_t first,
                        uint32_t count, uint32_t total, int threads, uint32_t row_begin,
                        uint32_t row_end, float *combined, float *normal, float *albedo,
                        uint32_t *sample_count, float *light, oracle_stats *stats) {
    if (!s || !rs || total == 0u || (uint64_t)first + count > total) return -1;
    const uint32_t seq = oracle_sequence_size_for_samples(total);
    float *table = (float *)malloc(sizeof(float) * 4u * 256u * seq);
    oracle_sobol_table(seq, table);
    kparams *k = (kparams *)malloc(sizeof(kparams));
    setup_params(s, rs, table, seq, k);
    if (row_end > k->window_height) row_end = k->window_height;
    job j;
    memset(&j, 0, sizeof(j));
    j.scene = s;
    j.k = k;
    j.first = first;
    j.count = count;
    j.row_begin = row_begin;
    j.row_end = row_end;
    j.next_row = row_begin;
    j.combined = combined;
    j.normal = normal;
    j.albedo = albedo;
    j.sample_count = sample_count;
    j.light = light;
    pthread_mutex_init(&j.lock, NULL);
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tids[256];
    for (int t = 0; t < threads; ++t) pthread_create(&tids[t], NULL, worker, &j);
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    pthread_mutex_destroy(&j.lock);
    if (stats) {
        stats->rays = j.rays;
        stats->node_visits = j.nodes;
        stats->triangle_tests = j.tris;
        stats->samples = (uint64_t)(row_end - row_begin) * k->window_width * count;
    }
    free(k);
    free(table);
    return 0;
}

int oracle_render(const oracle_scene *s, const psx_render_settings *rs, uint32_t first,
                  uint32_t count, uint32_t total, int threads, float *combined, float *normal,
                  float *albedo, uint32_t *sample_count, oracle_stats *stats) {
    const uint32_t h = rs->window_height ? rs->window_height : rs->height;
    return oracle_render_rows(s, rs, first, count, total, threads, 0u, h, combined, normal,
                              albedo, sample_count, stats);
}

void oracle_write_pass(uint32_t kind, uint32_t channels, uint32_t pixel_count, float exposure,
                       const float *combined, const float *normal, const float *albedo,
                       const uint32_t *sample_count, float *out) {
    for (uint32_t i = 0; i < pixel_count; ++i) {
        const float den = (float)(sample_count[i] > 1u ? sample_count[i] : 1u);
        float v[4] = {0, 0, 0, 0};
        switch (kind) {
        case PSX_PASS_COMBINED:
            v[0] = combined[4 * i] * (exposure / den);
            v[1] = combined[4 * i + 1] * (exposure / den);
            v[2] = combined[4 * i + 2] * (exposure / den);
            v[3] = fclamp(1.0f - combined[4 * i + 3] / den, 0.0f, 1.0f);
            break;
        case PSX_PASS_NORMAL:
        case PSX_PASS_DENOISING_NORMAL:
            for (int c = 0; c < 4; ++c) v[c] = normal[4 * i + c] / den;
            break;
        case PSX_PASS_ALBEDO:
        case PSX_PASS_DENOISING_ALBEDO:
            for (int c = 0; c < 4; ++c) v[c] = albedo[4 * i + c] / den;
            break;
        case PSX_PASS_SAMPLE_COUNT:
            v[0] = v[1] = v[2] = v[3] = (float)sample_count[i];
            break;
        default:
            break;
        }
        for (uint32_t c = 0; c < channels; ++c) out[i * channels + c] = v[c < 3 ? c : 3];
    }
}

/* Cycles' safe_divide_even_color (src/luisa/path_tracer_common.cpp:516-548):
 * channels with a zero denominator take the others' quotients */
static void safe_divide_even_color(const float *n, const float *d, float *out) {
    float x = d[0] != 0.0f ? n[0] / d[0] : 0.0f;
    float y = d[1] != 0.0f ? n[1] / d[1] : 0.0f;
    float z = d[2] != 0.0f ? n[2] / d[2] : 0.0f;
    if (d[0] == 0.0f) {
        if (d[1] == 0.0f) {
            x = z;
            y = z;
        } else if (d[2] == 0.0f) {
            x = y;
            z = y;
        } else {
            x = 0.5f * (y + z);
        }
    } else if (d[1] == 0.0f) {
        if (d[2] == 0.0f) {
            y = x;
            z = x;
        } else {
            y = 0.5f * (x + z);
        }
    } else if (d[2] == 0.0f) {
        z = 0.5f * (x + y);
    }
    out[0] = x;
    out[1] = y;
    out[2] = z;
}

void oracle_write_light_pass(uint32_t kind, uint32_t channels, uint32_t pixel_count, float exposure,
                             const float *light, const float *albedo, const uint32_t *sample_count,
                             float *out) {
    /* pass kind -> LightPassBuffer slot, and the colour it is divided by */
    int slot = -1, divisor = -2; /* divisor: -1 albedo, >= 0 a light slot, -2 none */
    int scale = 1;               /* 1: x exposure / N, 0: / N only */
    switch (kind) {
    case PSX_PASS_DIFFUSE_DIRECT: slot = 0; divisor = -1; break;
    case PSX_PASS_DIFFUSE_INDIRECT: slot = 1; divisor = -1; break;
    case PSX_PASS_GLOSSY_DIRECT: slot = 2; divisor = 10; break;
    case PSX_PASS_GLOSSY_INDIRECT: slot = 3; divisor = 10; break;
    case PSX_PASS_TRANSMISSION_DIRECT: slot = 4; divisor = 11; break;
    case PSX_PASS_TRANSMISSION_INDIRECT: slot = 5; divisor = 11; break;
    case PSX_PASS_VOLUME_DIRECT: slot = 6; break;
    case PSX_PASS_VOLUME_INDIRECT: slot = 7; break;
    case PSX_PASS_EMISSION: slot = 8; break;
    case PSX_PASS_ENVIRONMENT: slot = 9; break;
    case PSX_PASS_GLOSSY_COLOR: slot = 10; scale = 0; break;
    case PSX_PASS_TRANSMISSION_COLOR: slot = 11; scale = 0; break;
    default: break;
    }
    for (uint32_t i = 0; i < pixel_count; ++i) {
        float v[4] = {0, 0, 0, 0};
        if (slot >= 0) {
            const float *l = light + 4u * ORACLE_LIGHT_PASSES * i;
            const float *sum = l + 4 * slot;
            const float den = (float)(sample_count[i] > 1u ? sample_count[i] : 1u);
            if (divisor != -2) {
                const float *color = divisor == -1 ? albedo + 4 * i : l + 4 * divisor;
                safe_divide_even_color(sum, color, v);
                for (int c = 0; c < 3; ++c) v[c] *= exposure;
                v[3] = 1.0f;
            } else if (scale) {
                for (int c = 0; c < 4; ++c) v[c] = sum[c] * (exposure / den);
            } else {
                for (int c = 0; c < 4; ++c) v[c] = sum[c] / den;
            }
        }
        for (uint32_t c = 0; c < channels; ++c) out[i * channels + c] = v[c < 3 ? c : 3];
    }
}

int oracle_primary_hits(const oracle_scene *s, const psx_render_settings *rs,
                        uint32_t sample_index, uint32_t total, uint32_t *out) {
    const uint32_t seq = oracle_sequence_size_for_samples(total);
    float *table = (float *)malloc(sizeof(float) * 4u * 256u * seq);
    oracle_sobol_table(seq, table);
    kparams *k = (kparams *)malloc(sizeof(kparams));
    setup_params(s, rs, table, seq, k);
    for (uint32_t y = 0; y < k->window_height; ++y)
        for (uint32_t x = 0; x < k->window_width; ++x) {
            const camera_ray_t cam = camera_ray(k, x + k->window_x, y + k->window_y, sample_index);
            const hit_t h = closest_hit(s, &cam.ray, PSX_VISIBILITY_CAMERA, UINT32_MAX, UINT32_MAX,
                                        UINT32_MAX, UINT32_MAX, NULL);
            uint32_t *r = out + 4u * ((size_t)y * k->window_width + x);
            r[0] = h.valid ? h.object : UINT32_MAX;
            r[1] = h.valid ? h.cprim : UINT32_MAX;
            r[2] = h.valid ? fbits(h.t) : 0u;
            r[3] = h.valid ? 1u : 0u;
        }
    free(k);
    free(table);
    return 0;
}

int oracle_trace_events(const oracle_scene *s, const psx_render_settings *rs,
                        uint32_t sample_index, uint32_t total, uint32_t *out) {
    const uint32_t seq = oracle_sequence_size_for_samples(total);
    float *table = (float *)malloc(sizeof(float) * 4u * 256u * seq);
    oracle_sobol_table(seq, table);
    kparams *k = (kparams *)malloc(sizeof(kparams));
    setup_params(s, rs, table, seq, k);
    const size_t per = PSX_TRACE_EVENTS * PSX_TRACE_WORDS;
    for (uint32_t y = 0; y < k->window_height; ++y)
        for (uint32_t x = 0; x < k->window_width; ++x) {
            trace_log log;
            log.events = 0u;
            log.slots = NULL;
            log.records = out + per * ((size_t)y * k->window_width + x);
            memset(log.records, 0, sizeof(uint32_t) * per);
            path_out o;
            integrate(s, k, x + k->window_x, y + k->window_y, sample_index, &o, &log, NULL);
        }
    free(k);
    free(table);
    return 0;
}

int oracle_path_trace(const oracle_scene *s, const psx_render_settings *rs, uint32_t x, uint32_t y,
                      uint32_t sample_index, uint32_t total, float *out_slots, uint32_t *out_records) {
    const uint32_t seq = oracle_sequence_size_for_samples(total);
    float *table = (float *)malloc(sizeof(float) * 4u * 256u * seq);
    oracle_sobol_table(seq, table);
    kparams *k = (kparams *)malloc(sizeof(kparams));
    setup_params(s, rs, table, seq, k);
    uint32_t records[PSX_TRACE_EVENTS * PSX_TRACE_WORDS];
    memset(records, 0, sizeof(records));
    memset(out_slots, 0, sizeof(float) * 4u * PSX_PATH_TRACE_SLOTS);
    trace_log log;
    log.records = records;
    log.events = 0u;
    log.slots = out_slots;
    path_out o;
    integrate(s, k, x + k->window_x, y + k->window_y, sample_index, &o, &log, NULL);
    if (out_records) memcpy(out_records, records, sizeof(records));
    free(k);
    free(table);
    return 0;
}

int oracle_sampler_probe(const psx_render_settings *rs, uint32_t sample_index, uint32_t total,
                         uint32_t first_dimension, uint32_t dimension_count, uint32_t *out) {
    const uint32_t seq = oracle_sequence_size_for_samples(total);
    float *table = (float *)malloc(sizeof(float) * 4u * 256u * seq);
    oracle_sobol_table(seq, table);
    const uint32_t ww = rs->window_width ? rs->window_width : rs->width;
    const uint32_t wh = rs->window_height ? rs->window_height : rs->height;
    const uint32_t per = 1u + 4u * dimension_count;
    sampler smp = {table, seq};
    for (uint32_t y = 0; y < wh; ++y)
        for (uint32_t x = 0; x < ww; ++x) {
            const uint32_t fx = x + rs->window_x, fy = y + rs->window_y;
            const uint32_t cy = rs->height - 1u - fy;
            const uint32_t hash = oracle_pixel_hash(fx, cy, rs->seed);
            uint32_t *r = out + (size_t)per * ((size_t)y * ww + x);
            r[0] = hash;
            for (uint32_t d = 0; d < dimension_count; ++d) {
                const float *v = sample_4d(&smp, sample_index, hash, first_dimension + d);
                for (int c = 0; c < 4; ++c) r[1 + 4 * d + c] = fbits(v[c]);
            }
        }
    free(table);
    return 0;
}

/* ======================================================================= */
/* ray queries (SceneTraversalComponent::closest / closest_shadow,          */
/* src/luisa/path_kernel_scene_traversal.cpp:282-526; trace_shadow,         */
/* src/luisa/path_tracer_geometry.cpp:75-174)                               */
/* ======================================================================= */

int oracle_scene_intersect(const oracle_scene *s, const psx_ray_query *queries, uint32_t count,
                           psx_ray_hit *hits) {
    if (!s || (count && (!queries || !hits))) return -1;
    for (uint32_t i = 0; i < count; ++i) {
        const psx_ray_query *q = &queries[i];
        psx_ray_hit *h = &hits[i];
        memset(h, 0, sizeof(*h));
        ray_t r;
        r.o = load3(q->origin);
        r.d = load3(q->direction);
        r.tmin = q->tmin;
        r.tmax = q->tmax;
        if (q->kind == PSX_QUERY_SHADOW) {
            uint32_t passes = 0u;
            const v3 tr = trace_shadow(s, r, q->visibility, q->source_object, q->source_primitive,
                                       q->light_object, q->light_primitive, q->transparent_depth,
                                       q->transparent_max, NULL, &passes, NULL);
            h->transmittance[0] = tr.x;
            h->transmittance[1] = tr.y;
            h->transmittance[2] = tr.z;
            h->hit = !(tr.x > 0.0f || tr.y > 0.0f || tr.z > 0.0f);
            h->traversals = passes;
            h->object = h->primitive = PSX_INVALID_ID;
            h->t = q->tmax;
        } else {
            const hit_t b = closest_hit(s, &r, q->visibility, q->source_object, q->source_primitive,
                                        q->light_object, q->light_primitive, NULL);
            h->traversals = 1u;
            h->hit = b.valid ? 1u : 0u;
            h->t = b.valid ? b.t : q->tmax;
            h->u = b.valid ? b.u : 0.0f;
            h->v = b.valid ? b.v : 0.0f;
            h->object = b.valid ? b.object : PSX_INVALID_ID;
            h->primitive = b.valid ? b.cprim : PSX_INVALID_ID;
        }
    }
    return 0;
}

/* ======================================================================= */
/* single-function probes: the record layouts of psx_backend_probe          */
/* ======================================================================= */

static mat4 affine12(const float *a) {
    mat4 m;
    memset(&m, 0, sizeof(m));
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) m.m[4 * c + r] = a[3 * c + r];
    m.m[15] = 1.0f;
    return m;
}

/* ---- homogeneous volume transport ---------------------------------------
 * Restatement of the reference's analytic homogeneous-medium estimator
 * (HomogeneousVolumeTransport, src/luisa/homogeneous_volume_transport.cpp),
 * its scatter probability guide (src/luisa/volume_scatter_probability.cpp:
 * 20-98) and direct-light distance proposals (src/luisa/
 * volume_direct_sampling.cpp), in the reference's operation order.  Pinned by
 * records 0-45 of tests/test_luisa_homogeneous_volume.cpp:1052-1100. */
typedef struct {
    v3 sigma_t, sigma_s, emission;
    int has_extinction, has_scatter, has_emission;
} vol_coeffs;

static v3 vol_exp(v3 v) { return V3(expf(v.x), expf(v.y), expf(v.z)); }
static float vol_channel(v3 v, uint32_t c) { return c < 2u ? (c == 0u ? v.x : v.y) : v.z; }
static int vol_any_nonzero(v3 v) { return v.x != 0.0f || v.y != 0.0f || v.z != 0.0f; }
static v3 vol_safe_divide3(v3 n, v3 d) {
    return V3(d.x != 0.0f ? n.x / d.x : 0.0f, d.y != 0.0f ? n.y / d.y : 0.0f, d.z != 0.0f ? n.z / d.z : 0.0f);
}
static v3 vol_safe_divide1(v3 n, float d) { return d != 0.0f ? vdivs(n, d) : V3(0.0f, 0.0f, 0.0f); }
static float clampf_(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

/* transmittance (homogeneous_volume_transport.cpp:28-32) */
static v3 vol_transmittance(v3 sigma_t, float distance) { return vol_exp(vscale(vneg(sigma_t), distance)); }

/* emission_integral: (1 - e^-tau) / sigma_t, its Taylor form below 1e-5
 * optical depth, or the plain length without extinction (:33-57) */
static v3 vol_emission_integral(const vol_coeffs *c, float distance) {
    const v3 depth = vscale(c->sigma_t, distance);
    const float d[3] = {depth.x, depth.y, depth.z}, s[3] = {c->sigma_t.x, c->sigma_t.y, c->sigma_t.z};
    float a[3];
    for (int i = 0; i < 3; ++i)
        a[i] = d[i] > 1.0e-5f ? (1.0f - expf(-d[i])) / s[i] : distance * (1.0f - 0.5f * d[i]);
    return vmul(c->emission, c->has_extinction ? V3(a[0], a[1], a[2]) : V3(distance, distance, distance));
}

/* channel_pdf: |throughput * albedo| normalised, uniform when the sum fails
 * Cycles' (1 - sum) < 1 test (:58-74) */
static v3 vol_channel_pdf(v3 albedo, v3 throughput) {
    const v3 w = vmul(throughput, albedo);
    const v3 weights = V3(fabsf(w.x), fabsf(w.y), fabsf(w.z));
    const float sum = weights.x + weights.y + weights.z;
    const float third = 1.0f / 3.0f;
    return (1.0f - sum) < 1.0f ? vdivs(weights, sum) : V3(third, third, third);
}

typedef struct { uint32_t channel; v3 pdf; float random; int matched; } vol_channel_sample;

/* sample_channel: inverse CDF over the channel pdf, random rescaled (:75-113) */
static vol_channel_sample vol_sample_channel(v3 albedo, v3 throughput, float random) {
    vol_channel_sample r;
    r.pdf = vol_channel_pdf(albedo, throughput);
    const int first = random < r.pdf.x;
    const int second = !first && random < r.pdf.x + r.pdf.y;
    const int third = !first && !second && random < r.pdf.x + r.pdf.y + r.pdf.z;
    r.matched = first || second || third;
    r.channel = first ? 0u : (second ? 1u : 2u);
    const float lower = first ? 0.0f : (second ? r.pdf.x : r.pdf.x + r.pdf.y);
    const float selected = first ? r.pdf.x : (second ? r.pdf.y : r.pdf.z);
    r.random = r.matched ? (random - lower) / selected : random;
    return r;
}

/* bounded_exponential_sample: inverse CDF of rate * e^(-rate t) on
 * [minimum, maximum] (:114-133) */
static float vol_bexp_sample(float random, float rate, float minimum, float maximum) {
    const float attenuation = 1.0f - expf(rate * (minimum - maximum));
    const float safe_rate = rate != 0.0f ? rate : 1.0f;
    return clampf_(minimum - logf(1.0f - random * attenuation) / safe_rate, minimum, maximum);
}

/* bounded_exponential_pdf per channel rate (:134-152) */
static v3 vol_bexp_pdf(float distance, v3 rate, float minimum, float maximum) {
    const v3 neg = vneg(rate);
    const v3 attenuation = vsub(vol_exp(vscale(neg, minimum)), vol_exp(vscale(neg, maximum)));
    const v3 numerator = vmul(rate, vol_exp(vscale(neg, clampf_(distance, minimum, maximum))));
    return vol_safe_divide3(numerator, attenuation);
}

static int vol_scatters(const vol_coeffs *c) { return c->has_scatter && vol_any_nonzero(c->sigma_s); }

typedef struct {
    v3 transmittance, emission, throughput, scatter_probability, channel_pdf;
    float distance, event_pdf, scatter_random, reservoir_random;
    uint32_t channel;
    int scattered, active;
} vol_sample_t;

/* sample_with_probability (:201-341); sample() passes the unguided
 * probability 1 - T (:153-200) */
static vol_sample_t vol_sample_with_probability(const vol_coeffs *c, float distance, v3 throughput,
                                                float scatter_random, float channel_random, v3 probability,
                                                int terminate) {
    const v3 one = V3(1.0f, 1.0f, 1.0f);
    vol_sample_t r;
    r.transmittance = vol_transmittance(c->sigma_t, distance);
    r.emission = vmul(throughput, vol_emission_integral(c, distance));
    const int eligible = !terminate && vol_scatters(c) && distance > 0.0f;
    r.scatter_probability = eligible ? probability : V3(0.0f, 0.0f, 0.0f);
    const v3 albedo = vol_safe_divide3(c->sigma_s, c->sigma_t);
    const v3 ms_albedo = vmul(vmul(albedo, vsub(one, r.transmittance)), throughput);
    const vol_channel_sample ch = vol_sample_channel(vadd(ms_albedo, r.transmittance), throughput, channel_random);
    const float channel_sp = vol_channel(r.scatter_probability, ch.channel);
    r.scattered = eligible && scatter_random < channel_sp;
    const float rescaled = scatter_random / (r.scattered ? channel_sp : 1.0f);
    const float sampled = vol_bexp_sample(rescaled, vol_channel(c->sigma_t, ch.channel), 0.0f, distance);
    const float scatter_pdf =
        vdot(vmul(vol_bexp_pdf(sampled, c->sigma_t, 0.0f, distance), r.scatter_probability), ch.pdf);
    const v3 scatter_tp =
        vmul(throughput, vol_safe_divide1(vmul(c->sigma_s, vol_transmittance(c->sigma_t, sampled)), scatter_pdf));
    const float transmit_pdf = vdot(vsub(one, r.scatter_probability), ch.pdf);
    const v3 transmit_tp = vmul(throughput, vol_safe_divide1(r.transmittance, transmit_pdf));
    const float transmit_rescaled =
        (scatter_random - channel_sp) / ((!r.scattered && eligible) ? 1.0f - channel_sp : 1.0f);
    r.throughput = r.scattered ? scatter_tp : transmit_tp;
    r.channel_pdf = ch.pdf;
    r.distance = r.scattered ? sampled : distance;
    r.event_pdf = r.scattered ? scatter_pdf : transmit_pdf;
    r.scatter_random = r.scattered ? rescaled : transmit_rescaled;
    r.reservoir_random = ch.random;
    r.channel = ch.channel;
    r.active = c->has_extinction || c->has_scatter || c->has_emission;
    return r;
}

static vol_sample_t vol_sample(const vol_coeffs *c, float distance, v3 throughput, float scatter_random,
                               float channel_random, int terminate) {
    const v3 p = (!terminate && vol_scatters(c)) ? vsub(V3(1.0f, 1.0f, 1.0f), vol_transmittance(c->sigma_t, distance))
                                                 : V3(0.0f, 0.0f, 0.0f);
    return vol_sample_with_probability(c, distance, throughput, scatter_random, channel_random, p, terminate);
}

/* HomogeneousVolumeScatterProbability::evaluate
 * (volume_scatter_probability.cpp:20-98) */
static v3 vol_scatter_probability(const vol_coeffs *c, float distance, int terminate, v3 scattered_radiance,
                                  v3 transmitted_radiance, float majorant, int enabled) {
    const int attenuation_only = terminate || !c->has_scatter || !vol_any_nonzero(c->sigma_s);
    const v3 attenuation = vsub(V3(1.0f, 1.0f, 1.0f), vol_exp(vscale(vneg(c->sigma_t), distance)));
    const v3 radiance = vadd(transmitted_radiance, scattered_radiance);
    const int history_empty = !vol_any_nonzero(radiance);
    const v3 empty_p = V3(c->sigma_t.x > 0.0f ? 0.5f : 0.0f, c->sigma_t.y > 0.0f ? 0.5f : 0.0f,
                          c->sigma_t.z > 0.0f ? 0.5f : 0.0f);
    const float optical = 1.0f - expf(-majorant);
    const float scale = fmaxf(attenuation.x, fmaxf(attenuation.y, attenuation.z)) / (optical != 0.0f ? optical : 1.0f);
    const v3 h = vscale(vol_safe_divide3(scattered_radiance, radiance), scale);
    const v3 history_p = V3(clampf_(h.x, 0.0f, 1.0f), clampf_(h.y, 0.0f, 1.0f), clampf_(h.z, 0.0f, 1.0f));
    const v3 guided = history_empty ? empty_p : history_p;
    const v3 defensive = vadd(vscale(attenuation, 0.25f), vscale(guided, 0.75f));
    if (attenuation_only) return V3(0.0f, 0.0f, 0.0f);
    return enabled ? defensive : attenuation;
}

/* VolumeDirectSampling::prepare (volume_direct_sampling.cpp): MIS splits the
 * random number between distance (< 0.5) and equiangular sampling */
static void vol_direct_prepare(uint32_t requested, float random, int enabled, float *out) {
    const int use_mis = requested == 3u;
    const int choose_distance = random < 0.5f;
    const uint32_t method = use_mis ? (choose_distance ? 1u : 2u) : requested;
    const float prepared = use_mis ? (choose_distance ? random * 2.0f : (random - 0.5f) * 2.0f) : random;
    const int active = enabled && requested != 0u;
    out[0] = (float)(active ? method : 0u);
    out[1] = prepared;
    out[2] = active && use_mis ? 1.0f : 0.0f;
    out[3] = active ? 1.0f : 0.0f;
}

typedef struct {
    float delta, perpendicular, theta_min, theta_max, theta_range;
    int valid, uniform;
} vol_equiangular_geometry;

static vol_equiangular_geometry vol_equiangular(v3 origin, v3 direction, v3 light, float minimum, float maximum) {
    vol_equiangular_geometry g;
    const v3 offset = vsub(light, origin);
    g.delta = vdot(offset, direction);
    const v3 perp = vsub(offset, vscale(direction, g.delta));
    g.perpendicular = sqrtf(vdot(perp, perp));
    g.valid = minimum < maximum && g.perpendicular != 0.0f;
    const float safe = g.valid ? g.perpendicular : 1.0f;
    g.theta_min = atan2f(minimum - g.delta, safe);
    g.theta_max = atan2f(maximum - g.delta, safe);
    g.theta_range = g.theta_max - g.theta_min;
    g.uniform = g.valid && g.theta_range < 1.0e-6f;
    return g;
}

static float vol_uniform_pdf(float minimum, float maximum) {
    const float length = maximum - minimum;
    return length != 0.0f ? 1.0f / length : 0.0f;
}

static void vol_sample_equiangular(v3 origin, v3 direction, v3 light, float minimum, float maximum, float random,
                                   float *out) {
    const vol_equiangular_geometry g = vol_equiangular(origin, direction, light, minimum, maximum);
    const float uniform_distance = minimum + random * (maximum - minimum);
    const float angle = random * g.theta_max + (1.0f - random) * g.theta_min;
    const float offset = g.perpendicular * tanf(angle);
    const float sampled = clampf_(g.delta + offset, minimum, maximum);
    const float denominator = g.theta_range * (g.perpendicular * g.perpendicular + offset * offset);
    const float analytic = g.perpendicular / (denominator != 0.0f ? denominator : 1.0f);
    const float pdf = g.uniform ? vol_uniform_pdf(minimum, maximum) : analytic;
    out[0] = g.valid ? (g.uniform ? uniform_distance : sampled) : 0.0f;
    out[1] = g.valid ? pdf : 0.0f;
    out[2] = g.valid ? 1.0f : 0.0f;
}

static float vol_equiangular_pdf(v3 origin, v3 direction, v3 light, float minimum, float maximum, float distance) {
    const vol_equiangular_geometry g = vol_equiangular(origin, direction, light, minimum, maximum);
    const float offset = distance - g.delta;
    const float denominator = g.theta_range * (g.perpendicular * g.perpendicular + offset * offset);
    const float analytic = g.perpendicular / (denominator != 0.0f ? denominator : 1.0f);
    const float pdf = g.uniform ? vol_uniform_pdf(minimum, maximum) : analytic;
    return g.valid ? pdf : 0.0f;
}

static float vol_power_heuristic(float a, float b) {
    const float a2 = a * a, b2 = b * b, sum = a2 + b2;
    return sum > 0.0f ? a2 / sum : 0.0f;
}

static void store3(float *o, v3 v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }

/* HomogeneousVolumeTransport::sample_direct (:386-566) on a ray from the
 * origin along +z, after VolumeDirectSampling::prepare(method, technique
 * random, enabled); out as PSX_PROBE_VOLUME operation 13 */
static void vol_sample_direct(const vol_coeffs *c, float distance, v3 throughput, float scatter_random,
                              float reservoir_random, uint32_t requested, float technique_random, v3 light,
                              float minimum, float maximum, float *out) {
    float st[4];
    vol_direct_prepare(requested, technique_random, 1, st);
    const uint32_t method = (uint32_t)st[0];
    const int use_mis = st[2] != 0.0f, enabled = st[3] != 0.0f;
    const v3 origin = V3(0.0f, 0.0f, 0.0f), direction = V3(0.0f, 0.0f, 1.0f), one = V3(1.0f, 1.0f, 1.0f);
    const v3 segment_t = vol_transmittance(c->sigma_t, distance);
    const int scattered = enabled && vol_scatters(c) && distance > 0.0f && minimum < maximum;
    const v3 albedo = vol_safe_divide3(c->sigma_s, c->sigma_t);
    const v3 volume_albedo = vmul(vmul(albedo, vsub(one, segment_t)), throughput);
    const vol_channel_sample ch = vol_sample_channel(volume_albedo, throughput, reservoir_random);
    const float rate = vol_channel(c->sigma_t, ch.channel);
    const float distance_sample = vol_bexp_sample(scatter_random, rate, minimum, maximum);
    const float distance_sample_pdf = vdot(vol_bexp_pdf(distance_sample, c->sigma_t, minimum, maximum), ch.pdf);
    float eq[3];
    vol_sample_equiangular(origin, direction, light, minimum, maximum, st[1], eq);
    const int equiangular = method == 2u;
    const float sampled = equiangular ? eq[0] : distance_sample;
    const float distance_pdf =
        equiangular ? vdot(vol_bexp_pdf(eq[0], c->sigma_t, minimum, maximum), ch.pdf) : distance_sample_pdf;
    const float equiangular_pdf =
        equiangular ? eq[1] : vol_equiangular_pdf(origin, direction, light, minimum, maximum, distance_sample);
    const float selected = equiangular ? equiangular_pdf : distance_pdf;
    const float competing = equiangular ? distance_pdf : equiangular_pdf;
    const float mis = use_mis ? 2.0f * vol_power_heuristic(selected, competing) : 1.0f;
    const int valid = scattered && (!equiangular || eq[2] != 0.0f) && selected > 0.0f;
    const v3 sampled_t = vol_transmittance(c->sigma_t, sampled);
    const v3 direct_tp = vscale(vmul(throughput, vol_safe_divide1(vmul(c->sigma_s, sampled_t), selected)), mis);
    store3(out, valid ? sampled_t : one);
    store3(out + 3, valid ? direct_tp : V3(0.0f, 0.0f, 0.0f));
    store3(out + 6, ch.pdf);
    out[9] = valid ? sampled : 0.0f;
    out[10] = valid ? distance_pdf : 0.0f;
    out[11] = valid ? equiangular_pdf : 0.0f;
    out[12] = valid ? mis : 0.0f;
    out[13] = (float)ch.channel;
    out[14] = (float)method;
    out[15] = use_mis ? 1.0f : 0.0f;
    out[16] = valid ? 1.0f : 0.0f;
}

/* ---- light intervals (VolumeLightInterval, src/luisa/volume_light_interval.cpp) */
typedef struct { float minimum, maximum; int valid; } light_interval;

static float iv_min(float a, float b) { return b < a ? b : a; }
static float iv_max(float a, float b) { return b > a ? b : a; }

/* solve_quadratic (:53-120): the stable root pair, or the linear root */
static light_interval iv_quadratic(float a, float b, float c) {
    light_interval r;
    const int valid_linear = a == 0.0f && b != 0.0f;
    const float linear_root = -c / (b != 0.0f ? b : 1.0f);
    const float discriminant = b * b - 4.0f * a * c;
    const int valid_quadratic = a != 0.0f && discriminant > -1.0e-5f;
    const float root = sqrtf(iv_max(discriminant, 0.0f));
    const float temporary = -0.5f * (b + copysignf(root, b));
    const float r1 = temporary / (a != 0.0f ? a : 1.0f);
    const float r2 = temporary != 0.0f ? c / temporary : r1;
    r.minimum = valid_quadratic ? (r2 < r1 ? r2 : r1) : linear_root;
    r.maximum = valid_quadratic ? (r2 > r1 ? r2 : r1) : linear_root;
    r.valid = valid_linear || valid_quadratic;
    return r;
}

static light_interval iv_and(float lo0, float hi0, float lo1, float hi1, int valid) {
    light_interval r = {iv_max(lo0, lo1), iv_min(hi0, hi1), 0};
    r.valid = valid && r.minimum < r.maximum;
    return r;
}

static light_interval iv_plane(v3 normal, v3 plane_to_origin, v3 direction, float lo, float hi) {
    const float cosine = vdot(direction, normal);
    const float distance = -vdot(plane_to_origin, normal) / cosine;
    light_interval r = {cosine > 0.0f ? iv_max(lo, distance) : lo, cosine > 0.0f ? hi : iv_min(hi, distance), 0};
    r.valid = r.minimum < r.maximum;
    return r;
}

static light_interval iv_slab(float blo, float bhi, float origin, float direction, float lo, float hi) {
    const int moving = direction != 0.0f;
    const float safe = moving ? direction : 1.0f;
    const float a = (blo - origin) / safe, b = (bhi - origin) / safe;
    if (moving) return iv_and(lo, hi, iv_min(a, b), iv_max(a, b), 1);
    light_interval r = {lo, hi, origin >= blo && origin <= bhi && lo < hi};
    return r;
}

static light_interval iv_box(v3 blo, v3 bhi, v3 o, v3 d, float lo, float hi) {
    const light_interval x = iv_slab(blo.x, bhi.x, o.x, d.x, lo, hi);
    const light_interval y = iv_slab(blo.y, bhi.y, o.y, d.y, x.minimum, x.maximum);
    light_interval z = iv_slab(blo.z, bhi.z, o.z, d.z, y.minimum, y.maximum);
    z.valid = x.valid && y.valid && z.valid;
    return z;
}

static light_interval iv_elliptic_cylinder(v3 origin, v3 direction, float ru, float rv, float lo, float hi) {
    const int radii = ru > 0.0f && rv > 0.0f;
    const float su = radii ? ru : 1.0f, sv = radii ? rv : 1.0f;
    float ox = origin.x / su, oy = origin.y / sv;
    const float dx = direction.x / su, dy = direction.y / sv;
    const float a = dx * dx + dy * dy;
    float b = ox * dx + oy * dy;
    const int movable = a != 0.0f;
    const float middle = -b / (movable ? a : 1.0f);
    ox = ox + dx * middle;
    oy = oy + dy * middle;
    b = ox * dx + oy * dy;
    const float c = (ox * ox + oy * oy) - 1.0f;
    const light_interval q = iv_quadratic(a, 2.0f * b, c);
    light_interval r = iv_and(lo, hi, q.minimum + middle, q.maximum + middle, q.valid);
    r.valid = radii && movable && r.valid;
    return r;
}

/* intersect_cone (:325-440): the nappe around `axis`, or the plane when the
 * cone opens to a half space */
static light_interval iv_cone(v3 axis, v3 apex_to_origin, v3 direction, float cos2, float lo, float hi) {
    const light_interval plane = iv_plane(axis, apex_to_origin, direction, lo, hi);
    if (cos2 < 1.0e-4f) return plane;
    const float length2 = vdot(direction, direction);
    const float inverse_length = 1.0f / sqrtf(iv_max(length2, 1.0e-30f));
    const v3 nd = vscale(direction, inverse_length);
    const float ad = vdot(axis, nd), ao = vdot(axis, apex_to_origin);
    const float a = ad * ad - cos2;
    const float b = 2.0f * (ad * ao - cos2 * vdot(nd, apex_to_origin));
    const float c = ao * ao - cos2 * vdot(apex_to_origin, apex_to_origin);
    const light_interval q = iv_quadratic(a, b, c);
    const int min_valid = ao + q.minimum * ad > 0.0f, max_valid = ao + q.maximum * ad > 0.0f;
    const int neither = !max_valid, only_max = max_valid && !min_valid;
    const float hmin = only_max ? q.maximum : (neither ? 0.0f : q.minimum);
    const float hmax = only_max ? FLT_MAX : (neither ? q.minimum : q.maximum);
    return iv_and(lo, hi, hmin * inverse_length, hmax * inverse_length,
                  length2 > 0.0f && q.valid && (min_valid || max_valid));
}

/* VolumeLightInterval::spot (:441-512) */
static light_interval iv_spot(v3 origin, v3 direction, float lo, float hi, v3 center, v3 ax, v3 ay, v3 az, v3 scale,
                              float radius, float spot_angle) {
    const v3 cx = vscale(ax, scale.x), cy = vscale(ay, scale.y), cz = vscale(az, scale.z);
    const v3 adj_x = vcross(cy, cz);
    const float det = vdot(cx, adj_x);
    const int transform_valid = fabsf(det) > 1.0e-20f;
    const float inv_det = 1.0f / (transform_valid ? det : 1.0f);
    const v3 row_x = vscale(adj_x, inv_det), row_y = vscale(vcross(cz, cx), inv_det),
             row_z = vscale(vcross(cx, cy), inv_det);
    const float half_angle = 0.5f * spot_angle;
    const float tangent = tanf(half_angle);
    const float denominator = tangent * tangent * iv_min(scale.x * scale.x, scale.y * scale.y);
    const int geometry_valid = transform_valid && denominator > 0.0f;
    const float shift = radius * sqrtf(1.0f + scale.z * scale.z / (geometry_valid ? denominator : 1.0f));
    const v3 shifted = vsub(vadd(origin, vscale(vneg(az), shift)), center);
    const v3 lo3 = V3(vdot(row_x, shifted), vdot(row_y, shifted), vdot(row_z, shifted));
    const v3 ld = V3(vdot(row_x, direction), vdot(row_y, direction), vdot(row_z, direction));
    const float ch = cosf(half_angle);
    light_interval r = iv_cone(V3(0.0f, 0.0f, -1.0f), lo3, ld, ch * ch, lo, hi);
    r.valid = r.valid && geometry_valid;
    return r;
}

/* VolumeLightInterval::area (:513-647) */
static light_interval iv_area(v3 ray_origin, v3 direction, float lo, float hi, v3 center, v3 au, v3 av, v3 az,
                              float length_u, float length_v, float spread, int ellipse) {
    const v3 normal = vneg(az);
    const v3 origin = vsub(ray_origin, center);
    const v3 lorig = V3(vdot(origin, au), vdot(origin, av), vdot(origin, normal));
    const v3 ldir = V3(vdot(direction, au), vdot(direction, av), vdot(direction, normal));
    const float half_u = 0.5f * length_u, half_v = 0.5f * length_v;
    const float tangent = spread == 3.14159265358979323846f ? FLT_MAX : tanf(0.5f * iv_max(spread, 0.0f));
    light_interval region;
    if (tangent < 1.0e-5f) {
        region = ellipse ? iv_elliptic_cylinder(lorig, ldir, half_u, half_v, lo, hi)
                         : iv_box(V3(-half_u, -half_v, 0.0f), V3(half_u, half_v, FLT_MAX), lorig, ldir, lo, hi);
    } else {
        const float extent = 0.5f * (ellipse ? iv_max(length_u, length_v) : sqrtf(length_u * length_u + length_v * length_v));
        const v3 apex_to_origin = vadd(origin, vscale(normal, extent / tangent));
        region = iv_cone(normal, apex_to_origin, direction, 1.0f / (1.0f + tangent * tangent), lo, hi);
    }
    light_interval side = iv_plane(normal, origin, direction, region.minimum, region.maximum);
    side.valid = region.valid && side.valid;
    return side;
}

/* ---- phase functions (include/psycles/luisa/cycles_volume_phase.h) */
static float phase_clamp_anisotropy(float g) {
    const float s = g > 0.0f ? 1.0f : (g < 0.0f ? -1.0f : 0.0f);
    return s * fminf(fabsf(g), 1.0f - 1.0e-3f);
}

static float phase_hg_pdf(float cosine, float g) {
    float result = 0.07957747154594766788f;
    if (fabsf(g) >= 1.0e-3f) {
        const float factor = 1.0f + g * (g - 2.0f * cosine);
        result = (1.0f - g * g) / (4.0f * 3.14159265358979323846f * factor * sqrtf(fmaxf(factor, 0.0f)));
    }
    return result;
}

static float phase_rayleigh_pdf(float cosine) { return (0.1875f * 0.31830988618379067154f) * (1.0f + cosine * cosine); }

static float phase_hg_sample_cosine(float g, float random) {
    float cosine = 1.0f - 2.0f * random;
    if (fabsf(g) >= 1.0e-3f) {
        const float ratio = (1.0f - g * g) / (1.0f - g * cosine);
        cosine = (1.0f + g * g - ratio * ratio) / (2.0f * g);
    }
    return cosine;
}

/* Cycles' fast_inv_cbrtf (include/psycles/luisa/cycles_fast_math.h:229-243) */
static float phase_inverse_cube_root(float value) {
    int32_t bits;
    memcpy(&bits, &value, 4);
    const int32_t guess = 0x54a24242 - bits / 3;
    float r;
    memcpy(&r, &guess, 4);
    r = (2.0f / 3.0f) * r + 1.0f / (3.0f * r * r * value);
    r = (2.0f / 3.0f) * r + 1.0f / (3.0f * r * r * value);
    return r;
}

static float phase_rayleigh_sample_cosine(float random) {
    const float a = 2.0f - 4.0f * random;
    const float inverse_u = -phase_inverse_cube_root(sqrtf(1.0f + a * a) + a);
    return 1.0f / inverse_u - inverse_u;
}

static v3 phase_direction(v3 axis, float cosine, float azimuth_sample) {
    const float phi = 2.0f * 3.14159265358979323846f * azimuth_sample;
    const float sine = sqrtf(fmaxf(1.0f - cosine * cosine, 0.0f));
    v3 t, b;
    make_orthonormals(axis, &t, &b);
    return vadd(vadd(vscale(t, sine * cosf(phi)), vscale(b, sine * sinf(phi))), vscale(axis, cosine));
}

static vol_coeffs vol_block(const float *in) {
    vol_coeffs c;
    const uint32_t flags = (uint32_t)in[9];
    c.sigma_t = load3(in);
    c.sigma_s = load3(in + 3);
    c.emission = load3(in + 6);
    c.has_extinction = (flags & 1u) != 0u;
    c.has_scatter = (flags & 2u) != 0u;
    c.has_emission = (flags & 4u) != 0u;
    return c;
}


/* PSX_PROBE_VOLUME record (layouts in include/psx.h) */
static void volume_probe(const float *in, float *out) {
    switch ((uint32_t)in[0]) {
    case 0u: store3(out, vol_transmittance(load3(in + 1), in[4])); break;
    case 1u: { const vol_coeffs c = vol_block(in + 1); store3(out, vol_emission_integral(&c, in[11])); break; }
    case 2u: store3(out, vol_channel_pdf(load3(in + 1), load3(in + 4))); break;
    case 3u: {
        const vol_channel_sample s = vol_sample_channel(load3(in + 1), load3(in + 4), in[7]);
        out[0] = (float)s.channel; out[1] = s.random; out[2] = s.matched ? 1.0f : 0.0f;
        store3(out + 3, s.pdf);
        break;
    }
    case 4u: out[0] = vol_bexp_sample(in[1], in[2], in[3], in[4]); break;
    case 5u: store3(out, vol_bexp_pdf(in[1], load3(in + 2), in[5], in[6])); break;
    case 6u:
    case 7u: {
        const vol_coeffs c = vol_block(in + 1);
        const vol_sample_t h = (uint32_t)in[0] == 6u
            ? vol_sample(&c, in[11], load3(in + 12), in[15], in[16], in[17] != 0.0f)
            : vol_sample_with_probability(&c, in[11], load3(in + 12), in[15], in[16], load3(in + 18), in[17] != 0.0f);
        store3(out, h.transmittance);
        store3(out + 3, h.emission);
        store3(out + 6, h.throughput);
        store3(out + 9, h.scatter_probability);
        store3(out + 12, h.channel_pdf);
        out[15] = h.distance; out[16] = h.event_pdf; out[17] = h.scatter_random; out[18] = h.reservoir_random;
        out[19] = (float)h.channel; out[20] = h.scattered ? 1.0f : 0.0f; out[21] = h.active ? 1.0f : 0.0f;
        break;
    }
    case 8u: {
        const vol_coeffs c = vol_block(in + 1);
        store3(out, vol_scatter_probability(&c, in[11], in[12] != 0.0f, load3(in + 13), load3(in + 16), in[19],
                                            in[20] != 0.0f));
        break;
    }
    case 9u: vol_direct_prepare((uint32_t)in[1], in[2], in[3] != 0.0f, out); break;
    case 10u: vol_sample_equiangular(load3(in + 1), load3(in + 4), load3(in + 7), in[10], in[11], in[12], out); break;
    case 11u: out[0] = vol_equiangular_pdf(load3(in + 1), load3(in + 4), load3(in + 7), in[10], in[11], in[12]); break;
    case 12u: out[0] = vol_power_heuristic(in[1], in[2]); break;
    case 14u: {
        const light_interval r = iv_spot(load3(in + 1), load3(in + 4), in[7], in[8], load3(in + 9), load3(in + 12),
                                         load3(in + 15), load3(in + 18), load3(in + 21), in[24], in[25]);
        out[0] = r.minimum; out[1] = r.maximum; out[2] = r.valid ? 1.0f : 0.0f;
        break;
    }
    case 16u: {
        olamp l;
        memset(&l, 0, sizeof(l));
        l.type = PSX_LIGHT_SPOT;
        l.position = load3(in + 4);
        l.radius = in[7];
        l.axis_x = load3(in + 9);
        l.axis_y = load3(in + 12);
        l.axis_z = load3(in + 15);
        l.axis_scale = load3(in + 18);
        l.spot_angle = in[24];
        l.spot_smooth = in[25];
        l.flags = (in[8] != 0.0f ? LAMP_SPHERE : 0u) | (in[23] != 0.0f ? LAMP_NORMALIZE : 0u);
        const v2 rnd = {in[21], in[22]};
        const lamp_sample_t seg = sample_point_or_spot_ex(&l, load3(in + 1), V3(0, 0, 0), 1, rnd, 1);
        const lamp_sample_t pos = sample_point_or_spot_ex(&l, load3(in + 1), V3(0, 0, 0), 1, rnd, 0);
        out[0] = seg.valid ? 1.0f : 0.0f; out[1] = seg.eval; out[2] = seg.pdf;
        out[3] = pos.valid ? 1.0f : 0.0f; out[4] = pos.eval; out[5] = pos.pdf;
        break;
    }
    case 17u: {
        const float g = phase_clamp_anisotropy(in[1]);
        const float cosine = phase_hg_sample_cosine(g, in[6]);
        const v3 d = phase_direction(load3(in + 3), cosine, in[7]);
        out[0] = phase_hg_pdf(in[2], g);
        out[1] = g;
        store3(out + 2, d);
        out[5] = phase_hg_pdf(cosine, g);
        out[6] = phase_rayleigh_pdf(in[2]);
        const float rc = phase_rayleigh_sample_cosine(in[6]);
        store3(out + 7, phase_direction(load3(in + 3), rc, in[7]));
        out[10] = phase_rayleigh_pdf(rc);
        break;
    }
    case 15u: {
        const light_interval r = iv_area(load3(in + 1), load3(in + 4), in[7], in[8], load3(in + 9), load3(in + 12),
                                         load3(in + 15), load3(in + 18), in[21], in[22], in[23], in[24] != 0.0f);
        out[0] = r.minimum; out[1] = r.maximum; out[2] = r.valid ? 1.0f : 0.0f;
        break;
    }
    case 13u: {
        const vol_coeffs c = vol_block(in + 1);
        vol_sample_direct(&c, in[11], load3(in + 12), in[15], in[16], (uint32_t)in[17], in[18], load3(in + 19), in[22],
                          in[23], out);
        break;
    }
    default: break;
    }
}

uint32_t oracle_probe_input_width(uint32_t function) {
    static const uint32_t w[] = {6u, 1u, 21u, 47u, 14u, 15u, 13u, 59u, 21u, 32u};
    return function < sizeof(w) / sizeof(w[0]) ? w[function] : 0u;
}

uint32_t oracle_probe_output_width(uint32_t function) {
    static const uint32_t w[] = {3u, 1u, 4u, 4u, 10u, 3u, 4u, 4u, 10u, 24u};
    return function < sizeof(w) / sizeof(w[0]) ? w[function] : 0u;
}

int oracle_probe(uint32_t function, const float *inputs, uint32_t count, float *outputs) {
    const uint32_t wi = oracle_probe_input_width(function), wo = oracle_probe_output_width(function);
    if (!wi || (count && (!inputs || !outputs))) return -1;
    for (uint32_t i = 0; i < count; ++i) {
        const float *in = inputs + (size_t)wi * i;
        float *out = outputs + (size_t)wo * i;
        switch (function) {
        case PSX_PROBE_FILM_CLAMP: {
            const v3 c = film_clamp(load3(in), (uint32_t)in[3], in[4], in[5]);
            out[0] = c.x; out[1] = c.y; out[2] = c.z;
            break;
        }
        case PSX_PROBE_T_OFFSET:
            out[0] = bitsf(intersection_t_offset_bits(in[0]));
            break;
        case PSX_PROBE_SELF_EXCLUSION: {
            const v3 o = origin_with_self_exclusion(load3(in), load3(in + 3), load3(in + 6), load3(in + 9),
                                                    load3(in + 12), load3(in + 15), load3(in + 18));
            out[0] = o.x; out[1] = o.y; out[2] = o.z;
            out[3] = source_covers_origin(load3(in + 6), load3(in + 9), load3(in + 12), load3(in + 15),
                                          load3(in + 18)) ? 1.0f : 0.0f;
            break;
        }
        case PSX_PROBE_SHADOW_TERMINATOR:
        case PSX_PROBE_SURFACE_SHADOW_ORIGIN: {
            const mat4 o2w = affine12(in + 35);
            const int applied = in[14] != 0.0f;
            const v3 ng = load3(in + 6), dir = load3(in + 9);
            const v3 p0 = load3(in + 17), p1 = load3(in + 20), p2 = load3(in + 23);
            shadow_origin_t o;
            if (function == PSX_PROBE_SURFACE_SHADOW_ORIGIN) {
                const mat4 w2o = affine12(in + 47);
                o = shadow_origin_core(load3(in), load3(in + 3), ng, dir, in[12], in[13] != 0.0f, &o2w, &w2o,
                                       applied, in[15], in[16], p0, p1, p2, load3(in + 26), load3(in + 29),
                                       load3(in + 32));
            } else {
                o = shadow_terminator_origin(load3(in), load3(in + 3), ng, dir, in[12], in[13] != 0.0f, &o2w,
                                             applied, in[15], in[16], p0, p1, p2, load3(in + 26), load3(in + 29),
                                             load3(in + 32));
            }
            out[0] = o.position.x; out[1] = o.position.y; out[2] = o.position.z;
            out[3] = o.skip_self ? 1.0f : 0.0f;
            break;
        }
        case PSX_PROBE_TRIANGLE_SAMPLE: {
            const v2 rnd = {in[12], in[13]};
            const light_sample_t ls = sample_triangle(load3(in), load3(in + 3), load3(in + 6), load3(in + 9), rnd);
            out[0] = ls.position.x; out[1] = ls.position.y; out[2] = ls.position.z;
            out[3] = ls.direction.x; out[4] = ls.direction.y; out[5] = ls.direction.z;
            out[6] = ls.distance;
            out[7] = ls.pdf;
            out[8] = ls.valid ? 1.0f : 0.0f;
            out[9] = use_solid_angle(load3(in), load3(in + 3), load3(in + 6), load3(in + 9)) ? 1.0f : 0.0f;
            break;
        }
        case PSX_PROBE_TRIANGLE_PDF: {
            const tri_pdf_t p = triangle_directional_pdf(load3(in), load3(in + 3), load3(in + 6), load3(in + 9),
                                                         load3(in + 12));
            out[0] = p.value;
            out[1] = p.valid ? 1.0f : 0.0f;
            out[2] = use_solid_angle(load3(in), load3(in + 6), load3(in + 9), load3(in + 12)) ? 1.0f : 0.0f;
            break;
        }
        case PSX_PROBE_DIFFUSE_EVAL: {
            closure_rec c;
            memset(&c, 0, sizeof(c));
            c.kind = KIND_DIFFUSE;
            c.normal = load3(in);
            c.roughness = in[3];
            c.color = load3(in + 4);
            const v3 value = diffuse_intensity(&c, load3(in + 7), load3(in + 10));
            out[0] = value.x; out[1] = value.y; out[2] = value.z;
            out[3] = fmaxf(vdot(c.normal, load3(in + 10)), 0.0f) * INV_PI_F;
            break;
        }
        case PSX_PROBE_AREA_SAMPLE: {
            olamp l;
            memset(&l, 0, sizeof(l));
            l.type = PSX_LIGHT_AREA;
            l.position = load3(in);
            l.axis_x = load3(in + 3);
            l.axis_y = load3(in + 6);
            l.axis_z = load3(in + 9);
            l.size_u = in[12];
            l.size_v = in[13];
            l.spread = in[14];
            const uint32_t shape = (uint32_t)in[15];
            l.flags = ((shape & 1u) ? LAMP_ELLIPSE : 0u) | ((shape & 2u) ? LAMP_NORMALIZE : 0u) |
                      (l.spread >= 3.14159265358979f - 1.0e-6f ? LAMP_FULL_SPREAD : 0u);
            const v2 rnd = {in[19], in[20]};
            const lamp_sample_t s = sample_area_light(&l, load3(in + 16), rnd);
            out[0] = s.position.x; out[1] = s.position.y; out[2] = s.position.z;
            out[3] = s.direction.x; out[4] = s.direction.y; out[5] = s.direction.z;
            out[6] = s.distance;
            out[7] = s.pdf;
            out[8] = s.eval;
            out[9] = s.valid ? 1.0f : 0.0f;
            break;
        }
        case PSX_PROBE_VOLUME:
            memset(out, 0, sizeof(float) * wo);
            volume_probe(in, out);
            break;
        default:
            return -1;
        }
    }
    return 0;
}

/* One diffuse closure of colour `color` at a surface with shading and
 * geometric normal +z seen from +z: the records of the reference's closure
 * fixture (tests/test_luisa_cycles_closure.cpp:693-755) that this path
 * models, out[4 * r + c] for records r = 0..8 and 10: closure count, index,
 * Cycles type (2, diffuse) and sample weight; weight and validity; normal;
 * sampled index, type, sample weight and rescaled selection; weight and
 * validity; normal; pdf twice, label and validity; direction; evaluation;
 * (record 9, roughness and eta, is not modelled and stays zero); closure and
 * sample runtime flags and the closure counts for colours 0.5e-5 and 1e-5. */
static uint32_t diffuse_closure_count(float value) {
    psx_closure_instr ins;
    memset(&ins, 0, sizeof(ins));
    ins.op = PSX_CLOSURE_DIFFUSE;
    ins.color = 0u;
    ins.roughness = 3u;
    float params[4] = {value, value, value, 0.0f};
    omaterial m;
    memset(&m, 0, sizeof(m));
    m.instr = &ins;
    m.count = 1u;
    m.params = params;
    m.param_count = 4u;
    closure_set cs;
    material_closures(&m, V3(0, 0, 1), &cs);
    return cs.count;
}

int oracle_diffuse_closure_probe(const float color[3], float u_lobe, const float u_dir[2], float *out) {
    psx_closure_instr ins;
    memset(&ins, 0, sizeof(ins));
    ins.op = PSX_CLOSURE_DIFFUSE;
    ins.color = 0u;
    ins.roughness = 3u;
    float params[4] = {color[0], color[1], color[2], 0.0f};
    omaterial m;
    memset(&m, 0, sizeof(m));
    m.instr = &ins;
    m.count = 1u;
    m.params = params;
    m.param_count = 4u;
    closure_set cs;
    material_closures(&m, V3(0, 0, 1), &cs);
    memset(out, 0, sizeof(float) * 44u);
    const closure_rec *c = &cs.c[0];
    const float type = 2.0f;
    out[0] = (float)cs.count; out[1] = 0.0f; out[2] = type; out[3] = cs.count ? c->sample_weight : 0.0f;
    out[4] = c->weight.x; out[5] = c->weight.y; out[6] = c->weight.z; out[7] = cs.count ? 1.0f : 0.0f;
    out[8] = c->normal.x; out[9] = c->normal.y; out[10] = c->normal.z;
    surface_t sf;
    memset(&sf, 0, sizeof(sf));
    sf.ng = sf.ns = sf.incoming = V3(0, 0, 1);
    const v2 ud = {u_dir[0], u_dir[1]};
    const bsdf_sample_t bs = sample_closures(&cs, &sf, u_lobe, ud, LOBE_ALL);
    const closure_rec *sc = &cs.c[bs.selected < cs.count ? bs.selected : 0u];
    out[12] = (float)bs.selected; out[13] = type; out[14] = sc->sample_weight;
    out[15] = bs.rescaled;
    out[16] = sc->weight.x; out[17] = sc->weight.y; out[18] = sc->weight.z; out[19] = 1.0f;
    out[20] = sc->normal.x; out[21] = sc->normal.y; out[22] = sc->normal.z;
    out[24] = bs.eval.pdf; out[25] = bs.eval.pdf; out[26] = (float)label_from_events(bs.eval.events);
    out[27] = bs.valid ? 1.0f : 0.0f;
    out[28] = bs.wi.x; out[29] = bs.wi.y; out[30] = bs.wi.z;
    out[32] = bs.eval.f.x; out[33] = bs.eval.f.y; out[34] = bs.eval.f.z;
    out[40] = (float)surface_runtime_flags(&cs, 0);
    out[41] = (float)bs.runtime_flags;
    out[42] = (float)diffuse_closure_count(0.5e-5f);
    out[43] = (float)diffuse_closure_count(1.0e-5f);
    return 0;
}
