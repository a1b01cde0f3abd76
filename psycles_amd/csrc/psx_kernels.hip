// psx_kernels.hip - megakernel render path and probes (gfx950).
//
// The "plain" integrator: one lane per pixel of an 8x8 tile per wave runs
// whole paths, tracing each path's shadow rays inline, and accumulates the
// batch's samples in order in registers before one film read-modify-write
// (src/luisa/path_kernel_setup.cpp:903-1063).  The grid stays at the
// device's resident capacity and waves pull 8x8 tiles from an atomic cursor.
// Probes expose primary hits, discrete path events and sampler dimensions.
#include "psx_launch.h"
#include "psx_path.h"

namespace psx {
namespace {

constexpr uint32_t kBlock = 128;

struct PathResult {
    f3 radiance, normal, albedo;
    float transparency;
};

// Runs one path to completion.  A single traversal call site serves the
// continuation ray and every shadow query so the loop inlines once.
template <bool COUNT, bool TRACE>
__device__ PathResult integrate(const SceneView &s, const LampView &lv, const KernelParams &k, const float4 *sobol,
                                const float *filter, uint32_t fx, uint32_t fy, uint32_t sample, Stack &st,
                                uint32_t &rays, uint32_t &nodes, uint32_t &tris, uint32_t *trace,
                                float4 *slots = nullptr) {
    PathVars p;
    init_path(k, sobol, filter, fx, fy, sample, p);
    if (TRACE && slots) trace_globals(k, sobol, fx, fy, sample, p, slots);
    ShadowRay sr;
    sr.active = false;
    bool cont = (p.bits & PATH_DONE) == 0u;
    bool shadow = false;
    f3 transmittance = V(1.0f, 1.0f, 1.0f);
    uint32_t tdepth = 0u;
    bool first_hit = false;  // path traces: the shadow ray's first evaluated surface
    while (cont || shadow) {
        const Ray ray = shadow ? sr.ray : p.ray;
        const TraceQuery q = shadow ? sr.q : path_query(p);
        const Hit hit = closest_hit<COUNT>(s, ray, q, st, nodes, tris);
        ++rays;
        if (shadow) {
            float4 *se = TRACE ? shadow_slots(slots, sr.event) : nullptr;
            if (TRACE && se && hit.valid && !first_hit && tdepth < k.transparent_max) {
                // trace_shadow's first hit (src/luisa/path_tracer_geometry.cpp:136-143)
                first_hit = true;
                trace_slot(se, 5u, V(trace_identity(hit.object), trace_identity(hit.cprim), 0.0f));
                trace_slot(se, 6u, V(hit.t, hit.u, hit.v));
            }
            if (!shadow_hit(s, k.transparent_max, hit, sr.ray, tdepth, transmittance)) continue;
            if (TRACE && se) {
                if (!first_hit) {
                    trace_slot(se, 5u, V(-1.0f, -1.0f, -1.0f));
                    trace_slot(se, 6u, V(0.0f, 0.0f, 0.0f));
                }
                trace_slot(se, 7u, transmittance);
                float4 *te = event_slots(slots, sr.event);
                if (te && (transmittance.x > 0.0f || transmittance.y > 0.0f || transmittance.z > 0.0f))
                    trace_slot(te, 47u, sr.L * transmittance);
            }
            apply_shadow<false>(k, p, sr.L, sr.depth, transmittance, sr.dw, sr.gw);
            shadow = false;
            continue;
        }
        uint32_t *ev = (TRACE && p.step < 8u) ? trace + p.step * 8u : nullptr;
        cont = path_step<TRACE, false>(s, lv, k, sobol, p, hit, sr, ev, slots);
        if (sr.active) {
            shadow = true;
            transmittance = V(1.0f, 1.0f, 1.0f);
            tdepth = sr.tdepth;
            first_hit = false;
        }
    }
    PathResult out;
    out.radiance = p.radiance;
    out.normal = p.normal;
    out.albedo = p.albedo;
    out.transparency = p.transparency;
    return out;
}

// wave-level reduction then one atomic per wave
__device__ __forceinline__ void add_counter(unsigned long long *dst, uint32_t value) {
    unsigned long long v = value;
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset, 64);
    if ((threadIdx.x & 63) == 0 && v != 0ull) atomicAdd(dst, v);
}

__device__ __forceinline__ uint32_t tile_count(const KernelParams &k) {
    return ((k.window_width + 7u) >> 3) * ((k.window_height + 7u) >> 3);
}

// pixel of this lane in 8x8 tile `tile`
__device__ __forceinline__ bool tile_pixel(const KernelParams &k, uint32_t tile, uint32_t &px, uint32_t &py) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (k.window_width + 7u) >> 3;
    px = (tile % tiles_x) * 8u + (lane & 7u);
    py = (tile / tiles_x) * 8u + (lane >> 3);
    return px < k.window_width && py < k.window_height;
}

__device__ __forceinline__ Stack block_stack(uint32_t *lds, uint32_t *spill) {
    return make_stack(lds, spill, kBlock, threadIdx.x, blockIdx.x * kBlock + threadIdx.x, gridDim.x * kBlock);
}

// next 8x8 tile for this wave from the launch's tile cursor (one atomic per wave)
__device__ __forceinline__ uint32_t fetch_tile(uint32_t *cursor) {
    uint32_t tile = 0u;
    if ((threadIdx.x & 63u) == 0u) tile = atomicAdd(cursor, 1u);
    return __shfl(tile, 0, 64);
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void render_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                         const float *filter, FilmView film, uint32_t first,
                                                         uint32_t count, Counters *counters, uint32_t *spill,
                                                         uint32_t *cursor) {
    __shared__ uint32_t lds[kShortStack * kBlock];
    __shared__ float4 lamp_sphere[kLampCache];
    __shared__ uint32_t lamp_bits[kLampCache];
    const LampView lv = stage_lamps(s, lamp_sphere, lamp_bits);
    Stack st = block_stack(lds, spill);
    uint32_t rays = 0u, nodes = 0u, tris = 0u;
    const uint32_t tiles = tile_count(k);
    // waves pull tiles dynamically: path lengths vary a lot between tiles
    for (uint32_t tile = fetch_tile(cursor); tile < tiles; tile = fetch_tile(cursor)) {
        uint32_t px, py;
        if (!tile_pixel(k, tile, px, py)) continue;
        const size_t index = (size_t)py * k.window_width + px;
        float4 c = reinterpret_cast<float4 *>(film.combined)[index];
        float4 n = reinterpret_cast<float4 *>(film.normal)[index];
        float4 a = reinterpret_cast<float4 *>(film.albedo)[index];
        uint32_t completed = film.sample_count[index];
        for (uint32_t i = 0; i < count; ++i) {
            const PathResult r = integrate<COUNT, false>(s, lv, k, sobol, filter, px + k.window_x,
                                                         py + k.window_y, first + i, st, rays, nodes, tris, nullptr);
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

// discrete path events of one sample: 8 steps x 8 words per pixel
__global__ __launch_bounds__(kBlock) void trace_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                        const float *filter, uint32_t sample, uint32_t *records,
                                                        uint32_t *spill) {
    __shared__ uint32_t lds[kShortStack * kBlock];
    __shared__ float4 lamp_sphere[kLampCache];
    __shared__ uint32_t lamp_bits[kLampCache];
    const LampView lv = stage_lamps(s, lamp_sphere, lamp_bits);
    Stack st = block_stack(lds, spill);
    const uint32_t tiles = tile_count(k);
    const uint32_t waves = gridDim.x * (kBlock / 64u);
    for (uint32_t tile = blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6); tile < tiles; tile += waves) {
        uint32_t px, py;
        if (!tile_pixel(k, tile, px, py)) continue;
        uint32_t *rec = records + ((size_t)py * k.window_width + px) * 64u;
        for (int i = 0; i < 64; ++i) rec[i] = 0u;
        uint32_t rays = 0u, nodes = 0u, tris = 0u;
        integrate<false, true>(s, lv, k, sobol, filter, px + k.window_x, py + k.window_y, sample, st, rays, nodes,
                               tris, rec);
    }
}

// diagnostic path trace of one pixel and sample: the schema v2 slots and the
// pixel's discrete events; one lane of one block does the work
__global__ __launch_bounds__(kBlock) void path_trace_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                             const float *filter, uint32_t sample, uint32_t px,
                                                             uint32_t py, float4 *slots, uint32_t *records,
                                                             uint32_t *spill) {
    __shared__ uint32_t lds[kShortStack * kBlock];
    __shared__ float4 lamp_sphere[kLampCache];
    __shared__ uint32_t lamp_bits[kLampCache];
    const LampView lv = stage_lamps(s, lamp_sphere, lamp_bits);
    Stack st = block_stack(lds, spill);
    if (threadIdx.x != 0u) return;
    for (uint32_t i = 0; i < kTraceSlotCount; ++i) slots[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int i = 0; i < 64; ++i) records[i] = 0u;
    uint32_t rays = 0u, nodes = 0u, tris = 0u;
    integrate<false, true>(s, lv, k, sobol, filter, px + k.window_x, py + k.window_y, sample, st, rays, nodes, tris,
                           records, slots);
}

// primary hit identity: object, Cycles primitive, t bits, valid
__global__ __launch_bounds__(kBlock) void primary_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                          const float *filter, uint32_t sample, uint32_t *records,
                                                          uint32_t *spill) {
    __shared__ uint32_t lds[kShortStack * kBlock];
    Stack st = block_stack(lds, spill);
    const uint32_t tiles = tile_count(k);
    const uint32_t waves = gridDim.x * (kBlock / 64u);
    for (uint32_t tile = blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6); tile < tiles; tile += waves) {
        uint32_t px, py;
        if (!tile_pixel(k, tile, px, py)) continue;
        uint32_t hash;
        const Ray ray = camera_ray(k, sobol, filter, px + k.window_x, py + k.window_y, sample, hash);
        TraceQuery q;
        q.visibility = VIS_CAMERA;
        q.src_obj = q.src_prim = q.light_obj = q.light_prim = 0xffffffffu;
        uint32_t nodes = 0u, tris = 0u;
        const Hit h = closest_hit<false>(s, ray, q, st, nodes, tris);
        uint32_t *r = records + 4u * ((size_t)py * k.window_width + px);
        r[0] = h.valid ? h.object : 0xffffffffu;
        r[1] = h.valid ? h.cprim : 0xffffffffu;
        r[2] = h.valid ? __float_as_uint(h.t) : 0u;
        r[3] = h.valid ? 1u : 0u;
    }
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

__global__ void clear_kernel(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void material_table_kernel(SceneView s, float4 *shadow, float4 *emission, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 sh, em;
    material_tables(s, i, sh, em);
    shadow[i] = sh;
    emission[i] = em;
}


// per instance: 1 when every material its triangles can take (the mesh's
// slot table and the instance's overrides; no slot at all means no material)
// has no transparent extinction, so any hit on it zeroes a shadow ray
__global__ void instance_opacity_kernel(SceneView s, uint32_t *flags, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Instance &inst = s.instances[i];
    const Mesh &mesh = s.meshes[inst.mesh];
    auto opaque = [&](uint32_t material) {
        if (material == kNone) return true;
        const float4 e = s.material_shadow[material];
        return !(fmaxf(e.x, fmaxf(e.y, e.z)) > 0.0f);
    };
    bool all = true;
    for (uint32_t o = 0; o < inst.override_count; ++o) all = all && opaque(s.overrides[inst.override_offset + o]);
    for (uint32_t k = 0; k < mesh.slot_count; ++k) all = all && opaque(s.slots[mesh.slot_offset + k]);
    flags[i] = all ? 1u : 0u;
}

uint32_t tile_grid(const KernelParams &k, const void *kernel) {
    const uint32_t tiles = ((k.window_width + 7u) >> 3) * ((k.window_height + 7u) >> 3);
    const uint32_t blocks = (tiles + (kBlock / 64u) - 1u) / (kBlock / 64u);
    const uint32_t cap = resident_blocks(kernel, kBlock);
    return blocks < cap ? blocks : cap;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers called by the host library
// ---------------------------------------------------------------------------
uint32_t megakernel_lanes() {
    uint32_t lanes = 0u;
    const void *kernels[] = {reinterpret_cast<const void *>(render_kernel<false>),
                             reinterpret_cast<const void *>(render_kernel<true>),
                             reinterpret_cast<const void *>(trace_kernel),
                             reinterpret_cast<const void *>(primary_kernel)};
    for (const void *kern : kernels) {
        const uint32_t l = resident_blocks(kern, kBlock) * kBlock;
        lanes = l > lanes ? l : lanes;
    }
    return lanes;
}

// `cursor` is one device word, cleared here before the launch
hipError_t launch_render(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                         const FilmView &film, uint32_t first, uint32_t count, Counters *counters, bool count_nodes,
                         uint32_t *spill, uint32_t *cursor, hipStream_t stream) {
    hipLaunchKernelGGL(clear_kernel, dim3(1), dim3(64), 0, stream, cursor, 1u);
    if (count_nodes) {
        const uint32_t blocks = tile_grid(k, reinterpret_cast<const void *>(render_kernel<true>));
        if (blocks == 0u) return hipSuccess;
        hipLaunchKernelGGL(render_kernel<true>, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, film,
                           first, count, counters, spill, cursor);
    } else {
        const uint32_t blocks = tile_grid(k, reinterpret_cast<const void *>(render_kernel<false>));
        if (blocks == 0u) return hipSuccess;
        hipLaunchKernelGGL(render_kernel<false>, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, film,
                           first, count, counters, spill, cursor);
    }
    return hipGetLastError();
}

hipError_t launch_trace(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                        uint32_t sample, uint32_t *records, uint32_t *spill, hipStream_t stream) {
    const uint32_t blocks = tile_grid(k, reinterpret_cast<const void *>(trace_kernel));
    if (blocks == 0u) return hipSuccess;
    hipLaunchKernelGGL(trace_kernel, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, sample, records,
                       spill);
    return hipGetLastError();
}

hipError_t launch_path_trace(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                             uint32_t sample, uint32_t px, uint32_t py, float4 *slots, uint32_t *records,
                             uint32_t *spill, hipStream_t stream) {
    hipLaunchKernelGGL(path_trace_kernel, dim3(1), dim3(kBlock), 0, stream, s, k, sobol, filter, sample, px, py,
                       slots, records, spill);
    return hipGetLastError();
}

hipError_t launch_primary(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                          uint32_t sample, uint32_t *records, uint32_t *spill, hipStream_t stream) {
    const uint32_t blocks = tile_grid(k, reinterpret_cast<const void *>(primary_kernel));
    if (blocks == 0u) return hipSuccess;
    hipLaunchKernelGGL(primary_kernel, dim3(blocks), dim3(kBlock), 0, stream, s, k, sobol, filter, sample, records,
                       spill);
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

hipError_t launch_material_tables(const SceneView &s, float4 *shadow, float4 *emission, uint32_t count,
                                  hipStream_t stream) {
    if (count == 0u) return hipSuccess;
    hipLaunchKernelGGL(material_table_kernel, dim3((count + 63u) / 64u), dim3(64), 0, stream, s, shadow, emission,
                       count);
    return hipGetLastError();
}

// after launch_material_tables on the same stream (reads s.material_shadow)
hipError_t launch_instance_opacity(const SceneView &s, uint32_t *flags, uint32_t count, hipStream_t stream) {
    if (count == 0u) return hipSuccess;
    hipLaunchKernelGGL(instance_opacity_kernel, dim3((count + 63u) / 64u), dim3(64), 0, stream, s, flags, count);
    return hipGetLastError();
}

}  // namespace psx
