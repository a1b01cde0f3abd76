// psx_wavefront.hip - wavefront integrator (the batched render path, gfx950).
//
// A dispatch of `count` samples holds count x padded_pixels paths in HBM and
// advances them in lock-step iterations of two kernels:
//
//   trace  persistent waves run closest-hit traversal for continuation rays,
//          or the whole transparent shadow loop for NEE shadow rays; lanes
//          whose query finished are refilled from the queue (one atomic per
//          refill) while the rest of the wave keeps traversing
//   shade  consumes the hit of each live path: pending shadow result, then one
//          path step (emission, RR, AOVs, NEE sampling, BSDF sampling); new
//          rays are appended to the next queue with one atomic per wave
//          (ballot + prefix count)
//
// Shadow results are applied at the start of the next shade, before that
// step's emission, which is the order in which the megakernel adds them, and a
// final pass adds each pixel's samples in sample order; the film is therefore
// bit-identical to the megakernel's.  Queue counts live in device memory, so
// the host enqueues all iterations without synchronising; an empty iteration
// costs two near-empty launches.
#include <algorithm>

#include "psx_launch.h"
#include "psx_rtc.h"
#include "psx_wavefront_kernels.h"

namespace psx {

// ---------------------------------------------------------------------------
// launchers called by the host library
// ---------------------------------------------------------------------------
uint32_t wavefront_lanes() {
    const uint32_t a = resident_blocks(reinterpret_cast<const void *>(wf_trace_kernel<false>), kTraceBlock);
    const uint32_t b = resident_blocks(reinterpret_cast<const void *>(wf_trace_kernel<true>), kTraceBlock);
    const uint32_t c = resident_blocks(reinterpret_cast<const void *>(wf_trace_kernel<false, true>), kTraceBlock);
    const uint32_t d = resident_blocks(reinterpret_cast<const void *>(wf_trace_kernel<true, true>), kTraceBlock);
    return std::max(std::max(a, b), std::max(c, d)) * kTraceBlock;
}

// Entries the next trace queue and shade list need beyond the live ones:
// one partly used reservation per array and resident wave.
uint32_t wavefront_queue_headroom() {
    return static_cast<uint32_t>(device_cus()) * kMaxWavesPerCu * kMaxQueueChunk;
}
uint32_t wavefront_list_headroom() { return static_cast<uint32_t>(device_cus()) * kMaxWavesPerCu * kMaxListChunk; }

// 64 x the tiles of (1 << tile_shift) x (64 >> tile_shift) pixels over the window
uint32_t wavefront_padded_pixels(const KernelParams &k, uint32_t tile_shift) {
    const uint32_t tw = 1u << tile_shift, th = 64u >> tile_shift;
    return ((k.window_width + tw - 1u) >> tile_shift) * ((k.window_height + th - 1u) / th) * 64u;
}

// stores `value` at `dst` in stream order (a replayed graph reads it)
__global__ void store_word_kernel(uint32_t *dst, uint32_t value) {
    if (threadIdx.x == 0u) *dst = value;
}

hipError_t launch_store_word(uint32_t *dst, uint32_t value, hipStream_t stream) {
    hipLaunchKernelGGL(store_word_kernel, dim3(1), dim3(64), 0, stream, dst, value);
    return hipGetLastError();
}

using ShadeFn = void (*)(SceneView, KernelParams, const float4 *, WavefrontView, uint32_t);

// The shade variant of a dispatch as generated programs name it.
uint32_t shade_variant(const SceneView &s, bool light_passes) {
    return (light_passes ? kProgramLightPasses : 0u) | (scene_has_lights(s) ? kProgramLights : 0u);
}

// The shade variant for a dispatch: light-pass sums or not, lights in the
// scene or not.
ShadeFn shade_kernel(bool lp, bool lights) {
    if (lp) return lights ? wf_shade_kernel<true, true> : wf_shade_kernel<true, false>;
    return lights ? wf_shade_kernel<false, true> : wf_shade_kernel<false, false>;
}

// One dispatch of samples [first, first + count), or of [*first_at,
// *first_at + count) when `first_at` is given.  When `events` is given, it
// holds 2 * (k.max_path_steps + 1) events recorded around every trace launch.
// The sequence holds no memset and no host synchronisation, so it can be
// captured into a graph as it is.  `generated`, when given, is the scene's
// generated shade kernel for this variant, launched in place of the
// compiled-in one.
hipError_t launch_wavefront(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                            const FilmView &film, const WavefrontView &w, uint32_t first, uint32_t count,
                            const uint32_t *first_at, Counters *counters, bool count_nodes, uint32_t *spill,
                            hipEvent_t *events, const GeneratedShade *generated, hipStream_t stream) {
    const uint32_t total = count * w.padded_pixels;
    if (total == 0u) return hipSuccess;
    if (total > w.capacity || w.refill == 0u || w.refill > 64u || w.trace_chunk == 0u || w.queue_chunk < 128u ||
        w.queue_chunk > kMaxQueueChunk || w.list_chunk < 64u || w.list_chunk > kMaxListChunk || w.tile_shift < 3u ||
        w.tile_shift > 6u || (w.segments != 1u && w.segments != 8u))
        return hipErrorInvalidValue;
    const uint32_t iterations = k.max_path_steps + 1u;
    hipLaunchKernelGGL(reset_kernel, dim3((iterations + 1u + 63u) / 64u), dim3(64), 0, stream, w.counters,
                       iterations + 1u);
    hipLaunchKernelGGL(generate_kernel, dim3((total + kShadeBlock - 1u) / kShadeBlock), dim3(kShadeBlock), 0, stream, k,
                       sobol, filter, w, first, count, first_at);
    using TraceFn = void (*)(SceneView, KernelParams, WavefrontView, uint32_t, Counters *, uint32_t *);
    const bool segmented = w.segments == 8u;
    const TraceFn trace_fn = count_nodes ? (segmented ? wf_trace_kernel<true, true> : wf_trace_kernel<true, false>)
                                         : (segmented ? wf_trace_kernel<false, true> : wf_trace_kernel<false, false>);
    const uint32_t trace_blocks = resident_blocks(reinterpret_cast<const void *>(trace_fn), kTraceBlock);
    const bool lp = w.lp != nullptr;
    const bool lights = scene_has_lights(s);
    const ShadeFn shade_fn = shade_kernel(lp, lights);
    const uint32_t shade_blocks = resident_blocks(reinterpret_cast<const void *>(shade_fn), kShadeBlock);
    for (uint32_t it = 0; it < iterations; ++it) {
        if (events) {
            const hipError_t e = hipEventRecord(events[2u * it], stream);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(trace_fn, dim3(trace_blocks), dim3(kTraceBlock), 0, stream, s, k, w, it, counters, spill);
        if (events) {
            const hipError_t e = hipEventRecord(events[2u * it + 1u], stream);
            if (e != hipSuccess) return e;
        }
        if (generated) {
            SceneView sv = s;
            KernelParams kv = k;
            const float4 *sb = sobol;
            WavefrontView wv = w;
            uint32_t iv = it;
            void *args[] = {&sv, &kv, &sb, &wv, &iv};
            const hipError_t e = hipModuleLaunchKernel(generated->fn, generated->blocks, 1, 1, kShadeBlock, 1, 1, 0,
                                                       stream, args, nullptr);
            if (e != hipSuccess) return e;
        } else {
            hipLaunchKernelGGL(shade_fn, dim3(shade_blocks), dim3(kShadeBlock), 0, stream, s, k, sobol, w, it);
        }
    }
    hipLaunchKernelGGL(accumulate_kernel, dim3((w.padded_pixels + 255u) / 256u), dim3(256), 0, stream, k, w, film,
                       count);
    return hipGetLastError();
}

// A server batch, launched in steps so that the host can interleave other
// batches between its iterations: begin (counters, camera rays of every
// item), any number of iterations, end (in-order accumulation per item).
// `items` (device copy, path bases ascending) share the scene view `s` and
// the integrator fields of `k`; `paths` and `pixels` are the batch totals of
// count * padded and of padded over the items.
hipError_t launch_batch_begin(const BatchItem *items, uint32_t n_items, uint32_t paths, uint32_t iterations,
                              const WavefrontView &w, hipStream_t stream) {
    if (paths == 0u || n_items == 0u || paths > w.capacity || (paths & 63u) != 0u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_kernel, dim3((iterations + 1u + 63u) / 64u), dim3(64), 0, stream, w.counters,
                       iterations + 1u);
    hipLaunchKernelGGL(generate_batch_kernel, dim3((paths + kShadeBlock - 1u) / kShadeBlock), dim3(kShadeBlock), 0,
                       stream, items, n_items, w, paths);
    return hipGetLastError();
}

// `share` in (0, 1]: the fraction of the device's resident blocks the two
// kernels launch with, which leaves the rest of the compute units to batches
// on other streams.
hipError_t launch_batch_iteration(const SceneView &s, const KernelParams &k, const float4 *sobol,
                                  const WavefrontView &w, uint32_t it, Counters *counters, uint32_t *spill,
                                  float share, const GeneratedShade *generated, hipStream_t stream) {
    static const uint32_t trace_blocks =
        resident_blocks(reinterpret_cast<const void *>(wf_trace_kernel<false>), kTraceBlock);
    const ShadeFn shade_fn = shade_kernel(false, scene_has_lights(s));
    static const uint32_t shade_blocks_lights =
        resident_blocks(reinterpret_cast<const void *>(shade_kernel(false, true)), kShadeBlock);
    static const uint32_t shade_blocks_plain =
        resident_blocks(reinterpret_cast<const void *>(shade_kernel(false, false)), kShadeBlock);
    const uint32_t shade_blocks = scene_has_lights(s) ? shade_blocks_lights : shade_blocks_plain;
    const float f = share > 0.0f && share < 1.0f ? share : 1.0f;
    const uint32_t tb = std::max(1u, static_cast<uint32_t>(static_cast<float>(trace_blocks) * f));
    const uint32_t sb = std::max(
        1u, static_cast<uint32_t>(static_cast<float>(generated ? generated->blocks : shade_blocks) * f));
    hipLaunchKernelGGL(wf_trace_kernel<false>, dim3(tb), dim3(kTraceBlock), 0, stream, s, k, w, it, counters,
                       spill);
    if (generated) {
        SceneView sv = s;
        KernelParams kv = k;
        const float4 *sb4 = sobol;
        WavefrontView wv = w;
        uint32_t iv = it;
        void *args[] = {&sv, &kv, &sb4, &wv, &iv};
        const hipError_t e =
            hipModuleLaunchKernel(generated->fn, sb, 1, 1, kShadeBlock, 1, 1, 0, stream, args, nullptr);
        if (e != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL(shade_fn, dim3(sb), dim3(kShadeBlock), 0, stream, s, k, sobol, w, it);
    }
    return hipGetLastError();
}

hipError_t launch_batch_end(const BatchItem *items, uint32_t n_items, uint32_t pixels, const WavefrontView &w,
                            hipStream_t stream) {
    if ((pixels & 63u) != 0u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(accumulate_batch_kernel, dim3((pixels + 255u) / 256u), dim3(256), 0, stream, items, n_items,
                       w, pixels);
    return hipGetLastError();
}

}  // namespace psx
