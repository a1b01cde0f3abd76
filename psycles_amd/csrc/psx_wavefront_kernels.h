// psx_wavefront_kernels.h - device code of the wavefront integrator (gfx950):
// path-state loads and stores, camera-ray generation, the persistent trace
// kernel, the shade step and in-order accumulation, for single sessions and
// for server batches.  Included by psx_wavefront.hip (the compiled-in kernels
// and their launchers) and by the generated surface programs that psx_rtc.cpp
// compiles at run time.
#pragma once

#include "psx_path.h"

namespace psx {
namespace {

constexpr uint32_t kTraceBlock = 256;
constexpr uint32_t kShadeBlock = 256;
constexpr uint32_t kSkip = 0xffffffffu;  // queue/list entry of a padding lane
constexpr uint32_t kMaxWavesPerCu = 32u;

// Pixel of entry `rem` of a dispatch's sample plane: tiles of
// (1 << shift) x (64 >> shift) pixels, one per wave, row-major over the window.
__device__ __forceinline__ void tile_pixel(uint32_t rem, uint32_t window_width, uint32_t shift, uint32_t &px,
                                           uint32_t &py) {
    const uint32_t tile = rem >> 6, lane = rem & 63u;
    const uint32_t tw = 1u << shift;
    const uint32_t tiles_x = (window_width + tw - 1u) >> shift;
    px = (tile % tiles_x) * tw + (lane & (tw - 1u));
    py = (tile / tiles_x) * (64u >> shift) + (lane >> shift);
}

__device__ __forceinline__ void add_counter(unsigned long long *dst, uint32_t value) {
    unsigned long long v = value;
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset, 64);
    if ((threadIdx.x & 63) == 0 && v != 0ull) atomicAdd(dst, v);
}

// The state a shade step can change: the sample identity never changes
// after generation, and the primary Normal / Albedo only while the path is
// at bounce 0 (`aov`).  The MIS normal is read only by light sampling and
// forward-hit MIS, so a scene without lights (MIS = false) neither loads nor
// stores it.
template <bool MIS>
__device__ __forceinline__ void store_step(const WavefrontView &w, uint32_t i, const PathVars &p, bool aov) {
    w.ray_o[i] = make_float4(p.ray.o.x, p.ray.o.y, p.ray.o.z, p.ray.tmin);
    w.ray_d[i] = make_float4(p.ray.d.x, p.ray.d.y, p.ray.d.z, p.ray.tmax);
    w.aux[i] = make_uint4(p.src_obj, p.src_prim, p.step, p.bits);
    w.throughput[i] = make_float4(p.throughput.x, p.throughput.y, p.throughput.z, p.prev_pdf);
    if (MIS) w.mis_normal[i] = make_float4(p.mis_normal.x, p.mis_normal.y, p.mis_normal.z, 0.0f);
    w.radiance[i] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, p.transparency);
    if (aov) {
        w.normal[i] = make_float4(p.normal.x, p.normal.y, p.normal.z, 0.0f);
        w.albedo[i] = make_float4(p.albedo.x, p.albedo.y, p.albedo.z, 0.0f);
    }
    w.state0[i] = make_uint4(p.ps.flag, p.ps.visibility, p.ps.bounce, p.ps.diffuse_bounce);
    w.state1[i] = make_uint4(p.ps.glossy_bounce, p.ps.transmission_bounce, p.ps.transparent_bounce, p.ps.rng_offset);
}

__device__ __forceinline__ void store_path(const WavefrontView &w, uint32_t i, const PathVars &p) {
    w.ray_o[i] = make_float4(p.ray.o.x, p.ray.o.y, p.ray.o.z, p.ray.tmin);
    w.ray_d[i] = make_float4(p.ray.d.x, p.ray.d.y, p.ray.d.z, p.ray.tmax);
    w.aux[i] = make_uint4(p.src_obj, p.src_prim, p.step, p.bits);
    w.throughput[i] = make_float4(p.throughput.x, p.throughput.y, p.throughput.z, p.prev_pdf);
    w.mis_normal[i] = make_float4(p.mis_normal.x, p.mis_normal.y, p.mis_normal.z, 0.0f);
    w.radiance[i] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, p.transparency);
    w.normal[i] = make_float4(p.normal.x, p.normal.y, p.normal.z, 0.0f);
    w.albedo[i] = make_float4(p.albedo.x, p.albedo.y, p.albedo.z, 0.0f);
    w.state0[i] = make_uint4(p.ps.flag, p.ps.visibility, p.ps.bounce, p.ps.diffuse_bounce);
    w.state1[i] = make_uint4(p.ps.glossy_bounce, p.ps.transmission_bounce, p.ps.transparent_bounce, p.ps.rng_offset);
    w.ident[i] = make_uint2(p.sample, p.rng_hash);
}

// light-pass sums of a path, field-major: field f of path i at f * capacity + i
__device__ __forceinline__ void store_lp(const WavefrontView &w, uint32_t i, const LightSums &lp) {
    const size_t n = w.capacity;
    const f3 v[10] = {lp.dd, lp.di, lp.gd, lp.gi, lp.td, lp.ti, lp.emission, lp.environment, lp.path_dw, lp.path_gw};
#pragma unroll
    for (int f = 0; f < 10; ++f) w.lp[f * n + i] = make_float4(v[f].x, v[f].y, v[f].z, 0.0f);
}

__device__ __forceinline__ LightSums load_lp(const WavefrontView &w, uint32_t i) {
    const size_t n = w.capacity;
    f3 v[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) {
        const float4 q = w.lp[f * n + i];
        v[f] = V(q.x, q.y, q.z);
    }
    LightSums lp;
    lp.dd = v[0];
    lp.di = v[1];
    lp.gd = v[2];
    lp.gi = v[3];
    lp.td = v[4];
    lp.ti = v[5];
    lp.emission = v[6];
    lp.environment = v[7];
    lp.path_dw = v[8];
    lp.path_gw = v[9];
    return lp;
}

// A path's state for one shade step.  The primary Normal / Albedo are read
// only while the path is at bounce 0, the only steps that change them, and
// the MIS normal only with MIS (see store_step).
template <bool MIS>
__device__ __forceinline__ PathVars load_path(const WavefrontView &w, uint32_t i) {
    PathVars p;
    const float4 o = w.ray_o[i], d = w.ray_d[i];
    p.ray.o = V(o.x, o.y, o.z);
    p.ray.tmin = o.w;
    p.ray.d = V(d.x, d.y, d.z);
    p.ray.tmax = d.w;
    const uint4 a = w.aux[i];
    p.src_obj = a.x;
    p.src_prim = a.y;
    p.step = a.z;
    p.bits = a.w;
    const float4 t = w.throughput[i];
    p.throughput = V(t.x, t.y, t.z);
    p.prev_pdf = t.w;
    if (MIS) {
        const float4 mn = w.mis_normal[i];
        p.mis_normal = V(mn.x, mn.y, mn.z);
    } else {
        p.mis_normal = V(0.0f, 0.0f, 0.0f);
    }
    const float4 r = w.radiance[i];
    p.radiance = V(r.x, r.y, r.z);
    p.transparency = r.w;
    const uint4 s0 = w.state0[i], s1 = w.state1[i];
    if (s0.z == 0u) {
        const float4 n = w.normal[i];
        p.normal = V(n.x, n.y, n.z);
        const float4 al = w.albedo[i];
        p.albedo = V(al.x, al.y, al.z);
    } else {
        p.normal = p.albedo = V(0.0f, 0.0f, 0.0f);
    }
    p.ps.flag = s0.x;
    p.ps.visibility = s0.y;
    p.ps.bounce = s0.z;
    p.ps.diffuse_bounce = s0.w;
    p.ps.glossy_bounce = s1.x;
    p.ps.transmission_bounce = s1.y;
    p.ps.transparent_bounce = s1.z;
    p.ps.rng_offset = s1.w;
    const uint2 id = w.ident[i];
    p.sample = id.x;
    p.rng_hash = id.y;
    return p;
}

__global__ void reset_kernel(WavefrontCounters *counters, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = WavefrontCounters{0u, 0u, 0u, 0u};
}

// camera rays for samples [first, first + count) of every window pixel;
// `first_at`, when given, holds `first` in device memory (a replayed graph
// renders whatever range the host stored there)
__global__ __launch_bounds__(kShadeBlock) void generate_kernel(KernelParams k, const float4 *sobol,
                                                                const float *filter, WavefrontView w, uint32_t first,
                                                                uint32_t count, const uint32_t *first_at) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = count * w.padded_pixels;
    if (first_at) first = *first_at;
    bool live = false;
    if (i < total) {
        const uint32_t sample = first + i / w.padded_pixels;
        uint32_t px, py;
        tile_pixel(i % w.padded_pixels, k.window_width, w.tile_shift, px, py);
        PathVars p;
        if (px < k.window_width && py < k.window_height) {
            init_path(k, sobol, filter, px + k.window_x, py + k.window_y, sample, p);
            live = (p.bits & PATH_DONE) == 0u;
        } else {
            p.bits = PATH_DONE;
            p.ray.o = p.ray.d = V(0.0f, 0.0f, 0.0f);
            p.ray.tmin = p.ray.tmax = 0.0f;
            p.src_obj = p.src_prim = 0xffffffffu;
            p.throughput = p.radiance = p.normal = p.albedo = p.mis_normal = V(0.0f, 0.0f, 0.0f);
            p.prev_pdf = p.transparency = 0.0f;
            p.ps = PathState{0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
            p.sample = sample;
            p.rng_hash = 0u;
            p.step = 0u;
        }
        store_path(w, i, p);
        if (w.lp) {
            LightSums z;
            z.dd = z.di = z.gd = z.gi = z.td = z.ti = z.emission = z.environment = z.path_dw = z.path_gw =
                V(0.0f, 0.0f, 0.0f);
            store_lp(w, i, z);
        }
    }
    // iteration 0 lists every path in order; padding lanes are skip markers
    if (i < total) {
        w.queue[0][i] = live ? (i << 1) : kSkip;
        w.list[0][i] = live ? i : kSkip;
    }
    if (i == 0u) {
        w.counters[0].trace_count = total;
        w.counters[0].shade_count = total;
    }
}

// Persistent trace kernel.  Each lane owns one query at a time; when at least
// w.refill lanes of a wave have finished, the idle lanes take entries from
// the wave's reserved chunk of the queue and start them while the others keep
// their traversal state (Aila & Laine, "Understanding the efficiency of ray
// traversal on GPUs", 2009: persistent threads with ray replacement).
template <bool COUNT, bool SEGMENTED = false>
__global__ __launch_bounds__(kTraceBlock, 4) void wf_trace_kernel(SceneView s, KernelParams k, WavefrontView w,
                                                               uint32_t it, Counters *counters, uint32_t *spill) {
    __shared__ uint32_t lds[kShortStack * kTraceBlock];
    Stack st = make_stack(lds, spill, kTraceBlock, threadIdx.x, blockIdx.x * kTraceBlock + threadIdx.x,
                          gridDim.x * kTraceBlock);
    WavefrontCounters *c = w.counters + it;
    const uint32_t count = c->trace_count;
    const uint32_t *queue = w.queue[it & 1u];
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t rays = 0u, nodes = 0u, tris = 0u;
    Traversal t;
    bool active = false, exhausted = false, shadow = false;
    uint32_t path = 0u, tdepth = 0u;
    f3 transmittance = V(1.0f, 1.0f, 1.0f);
    // The wave's reservation [res, res_end) of queue entries.  One atomic on
    // the shared cursor reserves w.trace_chunk entries, which the wave hands to its
    // idle lanes over several refills: a returning atomic on one address
    // stalls the whole wave, and claiming per refill made it the largest
    // cost per ray.  `exhausted`: no chunk is left to reserve.
    uint32_t res = 0u, res_end = 0u;
    const uint32_t chunk = w.trace_chunk, refill = w.refill;
    // SEGMENTED (w.segments = 8): the queue is cut into 8 contiguous
    // segments, each with its own cursor.  A wave starts on the segment of
    // its block's XCD (workgroups are dispatched to the 8 XCDs round-robin),
    // so each XCD's L2 sees the rays of one region of the queue, and moves to
    // the next segment when its own is drained.
    constexpr uint32_t segments = SEGMENTED ? 8u : 1u;
    uint32_t seg = blockIdx.x % segments, tried = 0u;
    for (;;) {
        if (!exhausted || res < res_end) {
            const unsigned long long idle = __ballot(!active);
            const uint32_t want = static_cast<uint32_t>(__popcll(idle));
            if (res == res_end && !exhausted && want > 0u) {
                const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<long long>(idle)) - 1);
                if (!SEGMENTED) {
                    uint32_t b = 0u;
                    if (lane == leader) b = atomicAdd(&c->trace_fetch, chunk);
                    b = static_cast<uint32_t>(
                        __builtin_amdgcn_readlane(static_cast<int>(b), static_cast<int>(leader)));
                    if (b < count) {
                        res = b;
                        res_end = count - b > chunk ? b + chunk : count;
                    }
                    exhausted = b >= count || count - b <= chunk;
                } else {
                    uint32_t b = 0u, e = 0u;
                    if (lane == leader) {
                        for (;;) {
                            const uint32_t begin = static_cast<uint32_t>(
                                static_cast<unsigned long long>(count) * seg / segments);
                            const uint32_t end = static_cast<uint32_t>(
                                static_cast<unsigned long long>(count) * (seg + 1u) / segments);
                            const uint32_t got = atomicAdd(&c->segment_fetch[seg], chunk);
                            if (got < end - begin) {
                                b = begin + got;
                                e = end - b > chunk ? b + chunk : end;
                                break;
                            }
                            if (++tried == segments) break;
                            seg = seg + 1u == segments ? 0u : seg + 1u;
                        }
                    }
                    const int l = static_cast<int>(leader);
                    res = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), l));
                    res_end = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(e), l));
                    seg = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(seg), l));
                    tried = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(tried), l));
                    exhausted = tried == segments;
                }
            }
            const uint32_t avail = res_end - res;
            const uint32_t rank = static_cast<uint32_t>(__popcll(idle & below));
            const uint32_t qi = res + rank;
            res += want < avail ? want : avail;
            if (!active && rank < avail) {
                const uint32_t entry = queue[qi];
                if (entry != kSkip) {
                    path = entry >> 1;
                    shadow = (entry & 1u) != 0u;
                    if (shadow) {
                        const float4 o = w.shadow_o[path], d = w.shadow_d[path];
                        const uint4 sq = w.shadow_q[path];
                        t.ray.o = V(o.x, o.y, o.z);
                        t.ray.tmax = o.w;
                        t.ray.d = V(d.x, d.y, d.z);
                        t.ray.tmin = 0.0f;
                        t.q.visibility = VIS_SHADOW;
                        t.q.src_obj = sq.x;
                        t.q.src_prim = sq.y;
                        t.q.light_obj = sq.z;
                        t.q.light_prim = sq.w;
                        tdepth = __float_as_uint(w.shadow_l[path].w);
                        transmittance = V(1.0f, 1.0f, 1.0f);
                    } else {
                        const float4 o = w.ray_o[path], d = w.ray_d[path];
                        const uint4 a = w.aux[path];
                        t.ray.o = V(o.x, o.y, o.z);
                        t.ray.tmin = o.w;
                        t.ray.d = V(d.x, d.y, d.z);
                        t.ray.tmax = d.w;
                        t.q.visibility = contract_visibility(w.state0[path].y);
                        t.q.src_obj = a.x;
                        t.q.src_prim = a.y;
                        t.q.light_obj = t.q.light_prim = 0xffffffffu;
                    }
                    traversal_begin(s, t, st);
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (exhausted && res == res_end) break;
            continue;
        }
        // traverse until enough lanes are idle to refill (or all are, at the end)
        for (;;) {
            if (active && !traversal_step<COUNT>(s, t, st, nodes, tris)) {
                ++rays;
                if (!shadow) {
                    const Hit &h = t.best;
                    w.hit[path] = make_uint4(h.valid ? h.inst : kNone, h.prim, h.cprim, __float_as_uint(h.t));
                    w.hit_uv[path] = make_float2(h.u, h.v);
                    active = false;
                } else if (shadow_hit(s, k.transparent_max, t.best, t.ray, tdepth, transmittance)) {
                    w.shadow_t[path] = make_float4(transmittance.x, transmittance.y, transmittance.z, 0.0f);
                    active = false;
                } else {
                    // past a transparent surface: same shadow ray from the next float
                    traversal_begin(s, t, st);
                }
            }
            const uint32_t idle = static_cast<uint32_t>(__popcll(__ballot(!active)));
            if (idle == 64u || ((!exhausted || res < res_end) && idle >= refill)) break;
        }
    }
    add_counter(&counters->rays, rays);
    if (COUNT) {
        add_counter(&counters->nodes, nodes);
        add_counter(&counters->tris, tris);
    }
}

// The block's shade entries sorted by the shade class of their hit (miss
// first, then one class per closure structure), with skip entries last: a
// counting sort in LDS.  Paths of one class run the same closure code, so a
// wave diverges less; the order within a class is arbitrary, and no path's
// result depends on it.  Every thread of the block calls it.
__device__ __forceinline__ uint32_t sort_block(const SceneView &s, const WavefrontView &w, uint32_t entry,
                                               uint32_t *slot, uint32_t *counts) {
    uint32_t key = 0u;
    if (entry != kSkip) {
        const uint32_t inst = w.hit[entry].x;
        key = inst < s.instance_count ? s.instances[inst].shade_class : 0u;
        key = key < kShadeKeys ? key : kShadeKeys - 1u;
    }
    if (threadIdx.x < kShadeKeys) counts[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t rank = entry != kSkip ? atomicAdd(&counts[key], 1u) : 0u;
    __syncthreads();
    uint32_t offset = 0u, valid = 0u;
#pragma unroll
    for (uint32_t c = 0; c < kShadeKeys; ++c) {
        const uint32_t n = counts[c];
        offset += c < key ? n : 0u;
        valid += n;
    }
    if (entry != kSkip) slot[offset + rank] = entry;
    __syncthreads();
    const uint32_t out = threadIdx.x < valid ? slot[threadIdx.x] : kSkip;
    __syncthreads();
    return out;
}

// One shade launch: the body of wf_shade_kernel and of the shade kernels of
// generated surface programs.
template <bool LP, bool LIGHTS>
__device__ __forceinline__ void shade_launch(const SceneView &s, const KernelParams &k, const float4 *sobol,
                                             const WavefrontView &w, uint32_t it) {
    __shared__ float4 lamp_sphere[LIGHTS ? kLampCache : 1u];
    __shared__ uint32_t lamp_bits[LIGHTS ? kLampCache : 1u];
    LampView lv{lamp_sphere, lamp_bits, 0u};
    if (LIGHTS) lv = stage_lamps(s, lamp_sphere, lamp_bits);
    const WavefrontCounters *c = w.counters + it;
    WavefrontCounters *next = w.counters + it + 1u;
    const uint32_t count = c->shade_count;
    const uint32_t *list = w.list[it & 1u];
    uint32_t *next_queue = w.queue[(it + 1u) & 1u];
    uint32_t *next_list = w.list[(it + 1u) & 1u];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    // this wave's reservations [q_pos, q_end) in next_queue and [l_pos, l_end)
    // in next_list: a returning atomic on the shared counters per step would
    // stall the wave, as it did the trace kernel's refill
    uint32_t q_pos = 0u, q_end = 0u, l_pos = 0u, l_end = 0u;
    const uint32_t queue_chunk = w.queue_chunk, list_chunk = w.list_chunk;
    __shared__ uint32_t sort_slot[kShadeBlock];
    __shared__ uint32_t sort_count[kShadeKeys];
    const bool sorted = w.shade_sort != 0u;
    // the block steps over the list together (the sort is block-wide)
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        uint32_t entry = i < count ? list[i] : kSkip;
        if (sorted) entry = sort_block(s, w, entry, sort_slot, sort_count);
        bool cont = false, shadow = false;
        uint32_t path = 0u;
        if (entry != kSkip) {
            path = entry;
            PathVars p = load_path<LIGHTS>(w, path);
            const bool aov = p.ps.bounce == 0u;
            if (LP) p.lp = load_lp(w, path);
            if ((p.bits & PATH_SHADOW_PENDING) != 0u) {
                const float4 L = w.shadow_l[path];
                const float4 tr = w.shadow_t[path];
                f3 dw = V(0.0f, 0.0f, 0.0f), gw = dw;
                if (LP) {
                    const float4 a = w.shadow_split[path], b = w.shadow_split[w.capacity + path];
                    dw = V(a.x, a.y, a.z);
                    gw = V(b.x, b.y, b.z);
                }
                apply_shadow<LP>(k, p, V(L.x, L.y, L.z), __float_as_uint(w.shadow_d[path].w), V(tr.x, tr.y, tr.z),
                                 dw, gw);
                p.bits &= ~PATH_SHADOW_PENDING;
            }
            if ((p.bits & PATH_FINISHING) != 0u) {
                p.bits = (p.bits & ~PATH_FINISHING) | PATH_DONE;
            } else {
                const uint4 hv = w.hit[path];
                const float2 uv = w.hit_uv[path];
                Hit h;
                h.valid = hv.x != kNone;
                h.inst = hv.x;
                h.object = hv.x;
                h.prim = hv.y;
                h.cprim = hv.z;
                h.t = __uint_as_float(hv.w);
                h.u = uv.x;
                h.v = uv.y;
                ShadowRay sr;
                cont = path_step<false, LP, LIGHTS>(s, lv, k, sobol, p, h, sr, nullptr);
                shadow = sr.active;
                if (shadow) {
                    w.shadow_o[path] = make_float4(sr.ray.o.x, sr.ray.o.y, sr.ray.o.z, sr.ray.tmax);
                    w.shadow_d[path] = make_float4(sr.ray.d.x, sr.ray.d.y, sr.ray.d.z, __uint_as_float(sr.depth));
                    w.shadow_q[path] = make_uint4(sr.q.src_obj, sr.q.src_prim, sr.q.light_obj, sr.q.light_prim);
                    w.shadow_l[path] = make_float4(sr.L.x, sr.L.y, sr.L.z, __uint_as_float(sr.tdepth));
                    if (LP) {
                        w.shadow_split[path] = make_float4(sr.dw.x, sr.dw.y, sr.dw.z, 0.0f);
                        w.shadow_split[w.capacity + path] = make_float4(sr.gw.x, sr.gw.y, sr.gw.z, 0.0f);
                    }
                    p.bits |= PATH_SHADOW_PENDING;
                }
                if (!cont) p.bits |= shadow ? PATH_FINISHING : PATH_DONE;
            }
            store_step<LIGHTS>(w, path, p, aov);
            if (LP) store_lp(w, path, p.lp);
        }
        // the wave's continuation entries, then its shadow entries, into its
        // queue reservation, and its live paths into its list reservation;
        // an append that overflows a reservation continues in a new chunk
        const unsigned long long cm = __ballot(cont), sm = __ballot(shadow), lm = cm | sm;
        if (lm != 0ull) {
            const uint32_t nc = static_cast<uint32_t>(__popcll(cm)), ns = static_cast<uint32_t>(__popcll(sm));
            const uint32_t nq = nc + ns, nl = static_cast<uint32_t>(__popcll(lm));
            const uint32_t q_old = q_pos, q_avail = q_end - q_pos;
            uint32_t q_new = 0u;
            if (nq > q_avail) {
                if (lane == 0u) q_new = atomicAdd(&next->trace_count, queue_chunk);
                q_new = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(q_new)));
                q_pos = q_new + (nq - q_avail);
                q_end = q_new + queue_chunk;
            } else {
                q_pos += nq;
            }
            const uint32_t l_old = l_pos, l_avail = l_end - l_pos;
            uint32_t l_new = 0u;
            if (nl > l_avail) {
                if (lane == 0u) l_new = atomicAdd(&next->shade_count, list_chunk);
                l_new = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(l_new)));
                l_pos = l_new + (nl - l_avail);
                l_end = l_new + list_chunk;
            } else {
                l_pos += nl;
            }
            const uint32_t qr = cont ? static_cast<uint32_t>(__popcll(cm & below))
                                     : nc + static_cast<uint32_t>(__popcll(sm & below));
            const uint32_t qs = qr < q_avail ? q_old + qr : q_new + (qr - q_avail);
            if (cont) next_queue[qs] = path << 1;
            if (shadow && !cont) next_queue[qs] = (path << 1) | 1u;
            if (cont && shadow) {
                const uint32_t r2 = nc + static_cast<uint32_t>(__popcll(sm & below));
                next_queue[r2 < q_avail ? q_old + r2 : q_new + (r2 - q_avail)] = (path << 1) | 1u;
            }
            const uint32_t lr = static_cast<uint32_t>(__popcll(lm & below));
            if (cont || shadow) next_list[lr < l_avail ? l_old + lr : l_new + (lr - l_avail)] = path;
        }
    }
    // the unused ends of this wave's reservations become skip entries
    for (uint32_t j = q_pos + lane; j < q_end; j += 64u) next_queue[j] = kSkip;
    for (uint32_t j = l_pos + lane; j < l_end; j += 64u) next_list[j] = kSkip;
}

template <bool LP, bool LIGHTS>
__global__ __launch_bounds__(kShadeBlock, 2) void wf_shade_kernel(SceneView s, KernelParams k, const float4 *sobol,
                                                               WavefrontView w, uint32_t it) {
    shade_launch<LP, LIGHTS>(s, k, sobol, w, it);
}

// in-order film accumulation of the batch, one lane per window pixel
// adds one sample's light passes to a pixel's sums (w counts samples)
__device__ __forceinline__ void add_light_passes(float4 *sums, const WavefrontView &w, uint32_t path) {
    const size_t n = w.capacity;
    // LightSums fields in LightPassBuffer order; volume and colour passes stay zero
    const int field[kLightPasses] = {0, 1, 2, 3, 4, 5, -1, -1, 6, 7, -1, -1};
#pragma unroll
    for (uint32_t b = 0; b < kLightPasses; ++b) {
        float4 v = sums[b];
        if (field[b] >= 0) {
            const float4 q = w.lp[field[b] * n + path];
            v.x += q.x;
            v.y += q.y;
            v.z += q.z;
        } else {
            v.x += 0.0f;
            v.y += 0.0f;
            v.z += 0.0f;
        }
        v.w += 1.0f;
        sums[b] = v;
    }
}

__global__ __launch_bounds__(256) void accumulate_kernel(KernelParams k, WavefrontView w, FilmView film,
                                                         uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.padded_pixels) return;
    uint32_t px, py;
    tile_pixel(i, k.window_width, w.tile_shift, px, py);
    if (px >= k.window_width || py >= k.window_height) return;
    const size_t index = (size_t)py * k.window_width + px;
    float4 c = reinterpret_cast<float4 *>(film.combined)[index];
    float4 n = reinterpret_cast<float4 *>(film.normal)[index];
    float4 a = reinterpret_cast<float4 *>(film.albedo)[index];
    uint32_t completed = film.sample_count[index];
    float4 *light = film.light ? reinterpret_cast<float4 *>(film.light) + index * kLightPasses : nullptr;
    for (uint32_t s = 0; s < count; ++s) {
        const uint32_t path = s * w.padded_pixels + i;
        if (light) add_light_passes(light, w, path);
        const float4 r = w.radiance[path];
        const float4 pn = w.normal[path];
        const float4 pa = w.albedo[path];
        f3 rad = V(r.x, r.y, r.z);
        if (any_nan(rad)) rad = V(0.0f, 0.0f, 0.0f);
        c.x += rad.x;
        c.y += rad.y;
        c.z += rad.z;
        c.w += r.w;
        n.x += pn.x;
        n.y += pn.y;
        n.z += pn.z;
        n.w += 1.0f;
        a.x += pa.x;
        a.y += pa.y;
        a.z += pa.z;
        a.w += 1.0f;
        completed += 1u;
    }
    reinterpret_cast<float4 *>(film.combined)[index] = c;
    reinterpret_cast<float4 *>(film.normal)[index] = n;
    reinterpret_cast<float4 *>(film.albedo)[index] = a;
    film.sample_count[index] = completed;
}


// item of a batch-wide index: the last item whose base is <= i (bases rise)
template <bool PIXELS>
__device__ __forceinline__ uint32_t batch_item(const BatchItem *items, uint32_t n, uint32_t i) {
    uint32_t lo = 0u, len = n;
    while (len > 1u) {
        const uint32_t half = len >> 1;
        const uint32_t base = PIXELS ? items[lo + half].pixel_base : items[lo + half].path_base;
        if (base <= i) {
            lo += half;
            len -= half;
        } else {
            len = half;
        }
    }
    // every lane of a wave lies in the same item
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(lo)));
}

// camera rays for every item of a server batch
__global__ __launch_bounds__(kShadeBlock) void generate_batch_kernel(const BatchItem *items, uint32_t n_items,
                                                                      WavefrontView w, uint32_t total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) {
        w.counters[0].trace_count = total;
        w.counters[0].shade_count = total;
    }
    if (i - threadIdx.x % 64u >= total) return;
    const BatchItem &it = items[batch_item<false>(items, n_items, i)];
    const KernelParams &k = it.k;
    const uint32_t local = i - it.path_base;
    const uint32_t sample = it.first + local / it.padded;
    const uint32_t rem = local % it.padded;
    const uint32_t tile = rem >> 6, lane = rem & 63u;
    const uint32_t tiles_x = (k.window_width + 7u) >> 3;
    const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8u + (lane >> 3);
    PathVars p;
    bool live = false;
    if (px < k.window_width && py < k.window_height) {
        init_path(k, it.sobol, it.filter, px + k.window_x, py + k.window_y, sample, p);
        live = (p.bits & PATH_DONE) == 0u;
    } else {
        p.bits = PATH_DONE;
        p.ray.o = p.ray.d = V(0.0f, 0.0f, 0.0f);
        p.ray.tmin = p.ray.tmax = 0.0f;
        p.src_obj = p.src_prim = 0xffffffffu;
        p.throughput = p.radiance = p.normal = p.albedo = p.mis_normal = V(0.0f, 0.0f, 0.0f);
        p.prev_pdf = p.transparency = 0.0f;
        p.ps = PathState{0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
        p.sample = sample;
        p.rng_hash = 0u;
        p.step = 0u;
    }
    store_path(w, i, p);
    w.queue[0][i] = live ? (i << 1) : kSkip;
    w.list[0][i] = live ? i : kSkip;
}

// in-order film accumulation of every item of a server batch
__global__ __launch_bounds__(256) void accumulate_batch_kernel(const BatchItem *items, uint32_t n_items,
                                                               WavefrontView w, uint32_t total_pixels) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i - threadIdx.x % 64u >= total_pixels) return;
    const BatchItem &it = items[batch_item<true>(items, n_items, i)];
    const KernelParams &k = it.k;
    const uint32_t local = i - it.pixel_base;
    const uint32_t tile = local >> 6, lane = local & 63u;
    const uint32_t tiles_x = (k.window_width + 7u) >> 3;
    const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8u + (lane >> 3);
    if (px >= k.window_width || py >= k.window_height) return;
    const size_t index = (size_t)py * k.window_width + px;
    const FilmView &film = it.film;
    float4 c = reinterpret_cast<float4 *>(film.combined)[index];
    float4 n = reinterpret_cast<float4 *>(film.normal)[index];
    float4 a = reinterpret_cast<float4 *>(film.albedo)[index];
    uint32_t completed = film.sample_count[index];
    for (uint32_t s = 0; s < it.count; ++s) {
        const uint32_t path = it.path_base + s * it.padded + local;
        const float4 r = w.radiance[path];
        const float4 pn = w.normal[path];
        const float4 pa = w.albedo[path];
        f3 rad = V(r.x, r.y, r.z);
        if (any_nan(rad)) rad = V(0.0f, 0.0f, 0.0f);
        c.x += rad.x;
        c.y += rad.y;
        c.z += rad.z;
        c.w += r.w;
        n.x += pn.x;
        n.y += pn.y;
        n.z += pn.z;
        n.w += 1.0f;
        a.x += pa.x;
        a.y += pa.y;
        a.z += pa.z;
        a.w += 1.0f;
        completed += 1u;
    }
    reinterpret_cast<float4 *>(film.combined)[index] = c;
    reinterpret_cast<float4 *>(film.normal)[index] = n;
    reinterpret_cast<float4 *>(film.albedo)[index] = a;
    film.sample_count[index] = completed;
}

}  // namespace
}  // namespace psx
