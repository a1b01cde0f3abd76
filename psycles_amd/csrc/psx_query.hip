// psx_query.hip - ray queries and single-function probes (gfx950).
//
// intersect_kernel answers psx_ray_query records against a compiled scene
// with the render path's own traversal: closest hits under the Cycles order
// and the transparent shadow loop, whose any-hit traversal ends at the first
// candidate on an instance opaque to shadow rays.  Waves
// take 64 queries per atomic from a cursor; each lane keeps its stack in an
// LDS short stack with an HBM spill column, like the trace kernels.
//
// probe_kernel evaluates one device function per input record, so tests can
// hold the device code itself against the reference's function fixtures.
#include <mutex>

#include "../../include/psx.h"
#include "psx_launch.h"
#include "psx_path.h"
#include "psx_volume.h"

namespace psx {
namespace {

constexpr uint32_t kQueryBlock = 256;

__global__ void clear_words_kernel(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__device__ void run_query(const SceneView &s, const psx_ray_query &q, Stack &st, psx_ray_hit &h) {
    Ray r;
    r.o = V(q.origin[0], q.origin[1], q.origin[2]);
    r.d = V(q.direction[0], q.direction[1], q.direction[2]);
    r.tmin = q.tmin;
    r.tmax = q.tmax;
    TraceQuery tq;
    tq.visibility = q.visibility;
    tq.src_obj = q.source_object;
    tq.src_prim = q.source_primitive;
    tq.light_obj = q.light_object;
    tq.light_prim = q.light_primitive;
    uint32_t nodes = 0u, tris = 0u;
    h.u = h.v = 0.0f;
    h.reserved = h.reserved2 = 0u;
    if (q.kind == PSX_QUERY_SHADOW) {
        f3 transmittance = V(1.0f, 1.0f, 1.0f);
        uint32_t tdepth = q.transparent_depth, passes = 0u;
        for (;;) {
            bool opaque = false;
            const Hit b = closest_hit<false, true>(s, r, tq, st, nodes, tris, &opaque);
            ++passes;
            if (shadow_hit(s, q.transparent_max, b, r, tdepth, transmittance, opaque)) break;
        }
        h.t = q.tmax;
        h.object = h.primitive = PSX_INVALID_ID;
        h.hit = (transmittance.x > 0.0f || transmittance.y > 0.0f || transmittance.z > 0.0f) ? 0u : 1u;
        h.traversals = passes;
        h.transmittance[0] = transmittance.x;
        h.transmittance[1] = transmittance.y;
        h.transmittance[2] = transmittance.z;
    } else {
        const Hit b = closest_hit<false>(s, r, tq, st, nodes, tris);
        h.hit = b.valid ? 1u : 0u;
        h.t = b.valid ? b.t : q.tmax;
        h.u = b.valid ? b.u : 0.0f;
        h.v = b.valid ? b.v : 0.0f;
        h.object = b.valid ? b.object : PSX_INVALID_ID;
        h.primitive = b.valid ? b.cprim : PSX_INVALID_ID;
        h.traversals = 1u;
        h.transmittance[0] = h.transmittance[1] = h.transmittance[2] = 0.0f;
    }
}

// `cursor` (one word, zero at launch) hands out 64 queries per wave; `spill`
// holds kTraversalStackDepth words per lane of the grid.
__global__ __launch_bounds__(kQueryBlock) void intersect_kernel(SceneView s, const psx_ray_query *queries,
                                                                uint32_t count, psx_ray_hit *hits, uint32_t *cursor,
                                                                uint32_t *spill) {
    __shared__ uint32_t lds[kShortStack * kQueryBlock];
    Stack st = make_stack(lds, spill, kQueryBlock, threadIdx.x, blockIdx.x * kQueryBlock + threadIdx.x,
                          gridDim.x * kQueryBlock);
    const uint32_t lane = __lane_id();
    for (;;) {
        uint32_t base = 0u;
        if (lane == 0u) base = atomicAdd(cursor, 64u);
        base = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(base)));
        if (base >= count) break;
        const uint32_t i = base + lane;
        if (i < count) {
            const psx_ray_query q = queries[i];
            psx_ray_hit h;
            run_query(s, q, st, h);
            hits[i] = h;
        }
    }
}

__device__ __forceinline__ f3 in3(const float *p) { return V(p[0], p[1], p[2]); }

__device__ __forceinline__ Affine affine12(const float *p) {
    Affine a;
#pragma unroll
    for (int i = 0; i < 12; ++i) a.m[i] = p[i];
    return a;
}

__device__ __forceinline__ void out3(float *o, f3 v) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}

__device__ __forceinline__ VolumeCoefficients volume_block(const float *in) {
    const uint32_t flags = static_cast<uint32_t>(in[9]);
    return VolumeCoefficients{in3(in), in3(in + 3), in3(in + 6), (flags & 1u) != 0u, (flags & 2u) != 0u,
                              (flags & 4u) != 0u};
}

// PSX_PROBE_VOLUME: in[0] names the function, the rest its arguments
// (layouts in include/psx.h)
__device__ void volume_probe(const float *in, float *out) {
    switch (static_cast<uint32_t>(in[0])) {
    case 0u:
        out3(out, volume_transmittance(in3(in + 1), in[4]));
        break;
    case 1u:
        out3(out, volume_emission_integral(volume_block(in + 1), in[11]));
        break;
    case 2u:
        out3(out, volume_channel_pdf(in3(in + 1), in3(in + 4)));
        break;
    case 3u: {
        const VolumeChannelSample c = volume_sample_channel(in3(in + 1), in3(in + 4), in[7]);
        out[0] = static_cast<float>(c.channel);
        out[1] = c.random;
        out[2] = c.matched ? 1.0f : 0.0f;
        out3(out + 3, c.pdf);
        break;
    }
    case 4u:
        out[0] = volume_bounded_exponential_sample(in[1], in[2], in[3], in[4]);
        break;
    case 5u:
        out3(out, volume_bounded_exponential_pdf(in[1], in3(in + 2), in[5], in[6]));
        break;
    case 6u:
    case 7u: {
        const VolumeCoefficients c = volume_block(in + 1);
        const HomogeneousSample h =
            static_cast<uint32_t>(in[0]) == 6u
                ? volume_sample(c, in[11], in3(in + 12), in[15], in[16], in[17] != 0.0f)
                : volume_sample_with_probability(c, in[11], in3(in + 12), in[15], in[16], in3(in + 18),
                                                 in[17] != 0.0f);
        out3(out, h.transmittance);
        out3(out + 3, h.emission);
        out3(out + 6, h.throughput);
        out3(out + 9, h.scatter_probability);
        out3(out + 12, h.channel_pdf);
        out[15] = h.distance;
        out[16] = h.event_pdf;
        out[17] = h.scatter_random;
        out[18] = h.reservoir_random;
        out[19] = static_cast<float>(h.channel);
        out[20] = h.scattered ? 1.0f : 0.0f;
        out[21] = h.active ? 1.0f : 0.0f;
        break;
    }
    case 8u: {
        const ScatterGuide g{in3(in + 13), in3(in + 16), in[19], in[20] != 0.0f};
        out3(out, volume_scatter_probability(volume_block(in + 1), in[11], in[12] != 0.0f, g));
        break;
    }
    case 9u: {
        const DirectSamplingState d =
            volume_direct_prepare(static_cast<uint32_t>(in[1]), in[2], in[3] != 0.0f);
        out[0] = static_cast<float>(d.method);
        out[1] = d.random;
        out[2] = d.use_mis ? 1.0f : 0.0f;
        out[3] = d.enabled ? 1.0f : 0.0f;
        break;
    }
    case 10u: {
        const EquiangularSample e =
            volume_sample_equiangular(in3(in + 1), in3(in + 4), in3(in + 7), in[10], in[11], in[12]);
        out[0] = e.distance;
        out[1] = e.pdf;
        out[2] = e.valid ? 1.0f : 0.0f;
        break;
    }
    case 11u:
        out[0] = volume_equiangular_pdf(in3(in + 1), in3(in + 4), in3(in + 7), in[10], in[11], in[12]);
        break;
    case 12u:
        out[0] = volume_power_heuristic(in[1], in[2]);
        break;
    case 13u: {
        const DirectSamplingState st = volume_direct_prepare(static_cast<uint32_t>(in[17]), in[18], true);
        const HomogeneousDirectSample d =
            volume_sample_direct(volume_block(in + 1), in[11], in3(in + 12), in[15], in[16], st, V(0.0f, 0.0f, 0.0f),
                                 V(0.0f, 0.0f, 1.0f), in3(in + 19), in[22], in[23]);
        out3(out, d.transmittance);
        out3(out + 3, d.throughput);
        out3(out + 6, d.channel_pdf);
        out[9] = d.distance;
        out[10] = d.distance_pdf;
        out[11] = d.equiangular_pdf;
        out[12] = d.mis_weight;
        out[13] = static_cast<float>(d.channel);
        out[14] = static_cast<float>(d.method);
        out[15] = d.use_mis ? 1.0f : 0.0f;
        out[16] = d.scattered ? 1.0f : 0.0f;
        break;
    }
    case 14u: {
        const LightInterval r = volume_spot_interval(in3(in + 1), in3(in + 4), in[7], in[8], in3(in + 9), in3(in + 12),
                                                     in3(in + 15), in3(in + 18), in3(in + 21), in[24], in[25]);
        out[0] = r.minimum;
        out[1] = r.maximum;
        out[2] = r.valid ? 1.0f : 0.0f;
        break;
    }
    case 16u: {
        Lamp l{};
        l.type = kLampSpot;
        const f3 c = in3(in + 4), ax = in3(in + 9), ay = in3(in + 12), az = in3(in + 15), sc = in3(in + 18);
        const float cv[3] = {c.x, c.y, c.z}, xv[3] = {ax.x, ax.y, ax.z}, yv[3] = {ay.x, ay.y, ay.z},
                    zv[3] = {az.x, az.y, az.z}, sv[3] = {sc.x, sc.y, sc.z};
        for (int i = 0; i < 3; ++i) {
            l.position[i] = cv[i];
            l.axis_x[i] = xv[i];
            l.axis_y[i] = yv[i];
            l.axis_z[i] = zv[i];
            l.axis_scale[i] = sv[i];
        }
        l.radius = in[7];
        l.spot_angle = in[24];
        l.spot_smooth = in[25];
        l.flags = (in[8] != 0.0f ? kLampSphere : 0u) | (in[23] != 0.0f ? kLampNormalize : 0u);
        const f3 ref = in3(in + 1), zero = V(0.0f, 0.0f, 0.0f);
        const LampSample seg = sample_point_or_spot<true>(l, ref, zero, true, in[21], in[22]);
        const LampSample pos = sample_point_or_spot<false>(l, ref, zero, true, in[21], in[22]);
        out[0] = seg.valid ? 1.0f : 0.0f;
        out[1] = seg.eval;
        out[2] = seg.pdf;
        out[3] = pos.valid ? 1.0f : 0.0f;
        out[4] = pos.eval;
        out[5] = pos.pdf;
        break;
    }
    case 17u: {
        // Henyey-Greenstein phase of anisotropy in[1]: pdf at cosine in[2],
        // the clamped anisotropy, and a sample around axis in[3..5]
        const float g = phase_clamp_anisotropy(in[1]);
        const float cosine = phase_hg_sample_cosine(g, in[6]);
        const f3 d = phase_direction(in3(in + 3), cosine, in[7]);
        out[0] = phase_hg_pdf(in[2], g);
        out[1] = g;
        out3(out + 2, d);
        out[5] = phase_hg_pdf(cosine, g);
        out[6] = phase_rayleigh_pdf(in[2]);
        const float rc = phase_rayleigh_sample_cosine(in[6]);
        out3(out + 7, phase_direction(in3(in + 3), rc, in[7]));
        out[10] = phase_rayleigh_pdf(rc);
        break;
    }
    case 15u: {
        const LightInterval r = volume_area_interval(in3(in + 1), in3(in + 4), in[7], in[8], in3(in + 9), in3(in + 12),
                                                     in3(in + 15), in3(in + 18), in[21], in[22], in[23], in[24] != 0.0f);
        out[0] = r.minimum;
        out[1] = r.maximum;
        out[2] = r.valid ? 1.0f : 0.0f;
        break;
    }
    default:
        break;
    }
}

// record layouts of psx_backend_probe (include/psx.h)
__global__ void probe_kernel(uint32_t function, const float *inputs, uint32_t in_width, uint32_t count,
                             float *outputs, uint32_t out_width) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float *in = inputs + static_cast<size_t>(in_width) * i;
    float *out = outputs + static_cast<size_t>(out_width) * i;
    switch (function) {
    case PSX_PROBE_FILM_CLAMP:
        out3(out, film_clamp(in3(in), static_cast<uint32_t>(in[3]), in[4], in[5]));
        break;
    case PSX_PROBE_T_OFFSET:
        out[0] = next_float_up(in[0]);
        break;
    case PSX_PROBE_SELF_EXCLUSION: {
        const f3 oo = in3(in + 6), od = in3(in + 9), p0 = in3(in + 12), p1 = in3(in + 15), p2 = in3(in + 18);
        out3(out, origin_with_self_exclusion(in3(in), in3(in + 3), oo, od, p0, p1, p2));
        out[3] = source_covers_origin(oo, od, p0, p1, p2) ? 1.0f : 0.0f;
        break;
    }
    case PSX_PROBE_SHADOW_TERMINATOR:
    case PSX_PROBE_SURFACE_SHADOW_ORIGIN: {
        const Affine o2w = affine12(in + 35);
        const bool smooth = in[13] != 0.0f, applied = in[14] != 0.0f;
        ShadowOrigin o;
        if (function == PSX_PROBE_SURFACE_SHADOW_ORIGIN) {
            o = shadow_origin_core(in3(in), in3(in + 3), in3(in + 6), in3(in + 9), in[12], smooth, o2w,
                                   affine12(in + 47), applied, in[15], in[16], in3(in + 17), in3(in + 20),
                                   in3(in + 23), in3(in + 26), in3(in + 29), in3(in + 32));
        } else {
            o = shadow_terminator_origin(in3(in), in3(in + 3), in3(in + 6), in3(in + 9), in[12], smooth, o2w,
                                         applied, in[15], in[16], in3(in + 17), in3(in + 20), in3(in + 23),
                                         in3(in + 26), in3(in + 29), in3(in + 32));
        }
        out3(out, o.position);
        out[3] = o.skip_self ? 1.0f : 0.0f;
        break;
    }
    case PSX_PROBE_TRIANGLE_SAMPLE: {
        const f3 ref = in3(in), p0 = in3(in + 3), p1 = in3(in + 6), p2 = in3(in + 9);
        const LightSample ls = sample_triangle(ref, p0, p1, p2, in[12], in[13]);
        out3(out, ls.position);
        out3(out + 3, ls.direction);
        out[6] = ls.distance;
        out[7] = ls.pdf;
        out[8] = ls.valid ? 1.0f : 0.0f;
        out[9] = use_solid_angle(ref, p0, p1, p2) ? 1.0f : 0.0f;
        break;
    }
    case PSX_PROBE_TRIANGLE_PDF: {
        const f3 ref = in3(in), p0 = in3(in + 6), p1 = in3(in + 9), p2 = in3(in + 12);
        float value = 0.0f;
        const bool valid = triangle_pdf(ref, in3(in + 3), p0, p1, p2, value);
        out[0] = value;
        out[1] = valid ? 1.0f : 0.0f;
        out[2] = use_solid_angle(ref, p0, p1, p2) ? 1.0f : 0.0f;
        break;
    }
    case PSX_PROBE_DIFFUSE_EVAL: {
        const f3 n = in3(in);
        out3(out, diffuse_intensity(n, in[3], in3(in + 4), in3(in + 7), in3(in + 10)));
        out[3] = fmaxf(dot(n, in3(in + 10)), 0.0f) * INV_PI;
        break;
    }
    case PSX_PROBE_AREA_SAMPLE: {
        Lamp l{};
        l.type = kLampArea;
        const f3 c = in3(in), ax = in3(in + 3), ay = in3(in + 6), az = in3(in + 9);
        l.position[0] = c.x;
        l.position[1] = c.y;
        l.position[2] = c.z;
        l.axis_x[0] = ax.x;
        l.axis_x[1] = ax.y;
        l.axis_x[2] = ax.z;
        l.axis_y[0] = ay.x;
        l.axis_y[1] = ay.y;
        l.axis_y[2] = ay.z;
        l.axis_z[0] = az.x;
        l.axis_z[1] = az.y;
        l.axis_z[2] = az.z;
        l.size_u = in[12];
        l.size_v = in[13];
        l.spread = in[14];
        const uint32_t shape = static_cast<uint32_t>(in[15]);
        l.flags = ((shape & 1u) ? kLampEllipse : 0u) | ((shape & 2u) ? kLampNormalize : 0u) |
                  (l.spread >= 3.14159265358979f - 1.0e-6f ? kLampFullSpread : 0u);
        const LampSample s = sample_area_light(l, in3(in + 16), in[19], in[20]);
        out3(out, s.position);
        out3(out + 3, s.direction);
        out[6] = s.distance;
        out[7] = s.pdf;
        out[8] = s.eval;
        out[9] = s.valid ? 1.0f : 0.0f;
        break;
    }
    case PSX_PROBE_VOLUME:
        volume_probe(in, out);
        break;
    default:
        break;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free queries: the traversal scratch (the work cursor, padded to
// 256 bytes, then one spill column of kTraversalStackDepth words per lane for
// kNoAllocLanes lanes) lives in this file-scope device array, 16 MiB + 256 B
// on each device the library runs on.  Host code never passes the array's
// name to a kernel: setup_noalloc_scratch resolves its address on the current
// device once and keeps it per device.
// ---------------------------------------------------------------------------
constexpr uint32_t kNoAllocLanes = 65536u;
constexpr uint32_t kNoAllocCursorWords = 64u;
constexpr size_t kNoAllocWords = kNoAllocCursorWords + static_cast<size_t>(kNoAllocLanes) * kTraversalStackDepth;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mutex;
static uint32_t *g_noalloc_scratch[kMaxDevices] = {};  // the array's address per device

// ---------------------------------------------------------------------------
// launchers called by the host library
// ---------------------------------------------------------------------------
uint32_t query_resident_lanes() {
    return resident_blocks(reinterpret_cast<const void *>(intersect_kernel), kQueryBlock) * kQueryBlock;
}

// Lanes of a launch for `count` queries, at most `max_lanes` (a multiple of
// the block); the spill buffer must hold kTraversalStackDepth words per lane.
uint32_t query_lanes(uint32_t count, uint32_t max_lanes) {
    const uint32_t want = (count + kQueryBlock - 1u) / kQueryBlock * kQueryBlock;
    const uint32_t cap = max_lanes / kQueryBlock * kQueryBlock;
    return want < cap ? want : cap;
}

hipError_t launch_intersect(const SceneView &s, const psx_ray_query *queries, uint32_t count, psx_ray_hit *hits,
                            uint32_t *cursor, uint32_t *spill, uint32_t lanes, hipStream_t stream) {
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, stream, cursor, 1u);
    if (count == 0u || lanes < kQueryBlock) return hipGetLastError();
    hipLaunchKernelGGL(intersect_kernel, dim3(lanes / kQueryBlock), dim3(kQueryBlock), 0, stream, s, queries, count,
                       hits, cursor, spill);
    return hipGetLastError();
}

// Resolves the scratch array's address on the current device (once per
// device); called by psx_backend_create with its device current.
hipError_t setup_noalloc_scratch() {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_scratch[device]) return hipSuccess;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_scratch[device] = static_cast<uint32_t *>(ptr);
    return hipSuccess;
}

uint32_t noalloc_max_lanes() { return kNoAllocLanes; }

// The queries on the current device with the static scratch; nothing is
// allocated.  Fails with hipErrorNotInitialized before setup on this device.
hipError_t launch_intersect_noalloc(const SceneView &s, const psx_ray_query *queries, uint32_t count,
                                    psx_ray_hit *hits, hipStream_t stream) {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    uint32_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        if (device >= 0 && device < kMaxDevices) base = g_noalloc_scratch[device];
    }
    if (!base) return hipErrorNotInitialized;
    const uint32_t resident = query_resident_lanes();
    const uint32_t lanes = query_lanes(count, resident < kNoAllocLanes ? resident : kNoAllocLanes);
    return launch_intersect(s, queries, count, hits, base, base + kNoAllocCursorWords, lanes, stream);
}

hipError_t launch_probe(uint32_t function, const float *inputs, uint32_t in_width, uint32_t count, float *outputs,
                        uint32_t out_width, hipStream_t stream) {
    if (count == 0u) return hipSuccess;
    hipLaunchKernelGGL(probe_kernel, dim3((count + 255u) / 256u), dim3(256), 0, stream, function, inputs, in_width,
                       count, outputs, out_width);
    return hipGetLastError();
}

}  // namespace psx
