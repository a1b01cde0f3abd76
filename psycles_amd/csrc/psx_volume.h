// psx_volume.h - homogeneous volume transport primitives (device).
//
// The analytic homogeneous-medium estimator of Cycles as the reference's
// HomogeneousVolumeTransport states it (src/luisa/homogeneous_volume_transport.cpp):
// segment transmittance and emission, the throughput-weighted channel pick,
// the bounded free-flight distance sampler and its pdf, the
// scatter-versus-transmit decision with an optional guided scatter
// probability; the scatter probability guide
// (HomogeneousVolumeScatterProbability::evaluate,
// src/luisa/volume_scatter_probability.cpp:20-98); and the direct-light
// distance proposals (VolumeDirectSampling, src/luisa/volume_direct_sampling.cpp):
// technique preparation, equiangular sampling and pdf, the power heuristic.
//
// Each function keeps the reference's operation order, so the float results
// follow its arithmetic.  Include after psx_path.h (f3 and its operators).
#pragma once

namespace psx {

// volume_sample_* (include/psycles/luisa/volume_stack.h:19-25)
constexpr uint32_t kVolumeSampleNone = 0u, kVolumeSampleDistance = 1u, kVolumeSampleEquiangular = 2u,
                   kVolumeSampleMis = 3u;

struct VolumeCoefficients {
    f3 sigma_t, sigma_s, emission;
    bool has_extinction, has_scatter, has_emission;
};

struct VolumeChannelSample {
    uint32_t channel;
    f3 pdf;
    float random;
    bool matched;
};

struct HomogeneousSample {
    f3 transmittance, emission, throughput, scatter_probability, channel_pdf;
    float distance, event_pdf, scatter_random, reservoir_random;
    uint32_t channel;
    bool scattered, active;
};

struct ScatterGuide {
    f3 scattered_radiance, transmitted_radiance;
    float majorant_optical_depth;
    bool enabled;
};

struct DirectSamplingState {
    uint32_t method;
    float random;
    bool use_mis, enabled;
};

struct EquiangularSample {
    float distance, pdf;
    bool valid;
};

__device__ __forceinline__ f3 vol_exp(f3 v) { return V(expf(v.x), expf(v.y), expf(v.z)); }
__device__ __forceinline__ f3 vol_neg(f3 v) { return V(-v.x, -v.y, -v.z); }
__device__ __forceinline__ float vol_channel(f3 v, uint32_t c) { return c < 2u ? (c == 0u ? v.x : v.y) : v.z; }
__device__ __forceinline__ bool vol_any_nonzero(f3 v) { return v.x != 0.0f || v.y != 0.0f || v.z != 0.0f; }

// numerator / denominator per component, `fallback` where the denominator is 0
__device__ __forceinline__ f3 vol_safe_divide(f3 n, f3 d, float fallback = 0.0f) {
    return V(d.x != 0.0f ? n.x / d.x : fallback, d.y != 0.0f ? n.y / d.y : fallback,
             d.z != 0.0f ? n.z / d.z : fallback);
}
__device__ __forceinline__ f3 vol_safe_divide(f3 n, float d) {
    return d != 0.0f ? n / d : V(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ f3 volume_transmittance(f3 sigma_t, float distance) {
    return vol_exp(vol_neg(sigma_t) * distance);
}

__device__ inline f3 volume_emission_integral(const VolumeCoefficients &c, float distance) {
    const f3 depth = c.sigma_t * distance;
    f3 attenuated;
    const float d[3] = {depth.x, depth.y, depth.z}, s[3] = {c.sigma_t.x, c.sigma_t.y, c.sigma_t.z};
    float a[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const bool analytic = d[i] > 1.0e-5f;
        a[i] = analytic ? (1.0f - expf(-d[i])) / s[i] : distance * (1.0f - 0.5f * d[i]);
    }
    attenuated = V(a[0], a[1], a[2]);
    return c.emission * (c.has_extinction ? attenuated : V(distance, distance, distance));
}

__device__ inline f3 volume_channel_pdf(f3 albedo, f3 throughput) {
    const f3 w = throughput * albedo;
    const f3 weights = V(fabsf(w.x), fabsf(w.y), fabsf(w.z));
    const float sum = weights.x + weights.y + weights.z;
    // (1 - sum) < 1 rather than sum > 0: denormal sums count as empty
    const bool usable = (1.0f - sum) < 1.0f;
    const float third = 1.0f / 3.0f;
    return usable ? weights / sum : V(third, third, third);
}

__device__ inline VolumeChannelSample volume_sample_channel(f3 albedo, f3 throughput, float random) {
    const f3 pdf = volume_channel_pdf(albedo, throughput);
    const bool first = random < pdf.x;
    const bool second = !first && random < pdf.x + pdf.y;
    const bool third = !first && !second && random < pdf.x + pdf.y + pdf.z;
    const bool matched = first || second || third;
    const uint32_t channel = (first || second) ? (first ? 0u : 1u) : 2u;
    const float lower = (first || second) ? (first ? 0.0f : pdf.x) : pdf.x + pdf.y;
    const float selected = (first || second) ? (first ? pdf.x : pdf.y) : pdf.z;
    const float rescaled = (random - lower) / (matched ? selected : 1.0f);
    return VolumeChannelSample{channel, pdf, matched ? rescaled : random, matched};
}

__device__ inline float volume_bounded_exponential_sample(float random, float rate, float minimum, float maximum) {
    const float attenuation = 1.0f - expf(rate * (minimum - maximum));
    const float safe_rate = rate != 0.0f ? rate : 1.0f;
    return fminf(fmaxf(minimum - logf(1.0f - random * attenuation) / safe_rate, minimum), maximum);
}

__device__ inline f3 volume_bounded_exponential_pdf(float distance, f3 rate, float minimum, float maximum) {
    const f3 neg = vol_neg(rate);
    const f3 attenuation = vol_exp(neg * minimum) - vol_exp(neg * maximum);
    const float t = fminf(fmaxf(distance, minimum), maximum);
    const f3 numerator = rate * vol_exp(neg * t);
    return vol_safe_divide(numerator, attenuation);
}

__device__ __forceinline__ bool volume_scatters(const VolumeCoefficients &c) {
    return c.has_scatter && vol_any_nonzero(c.sigma_s);
}

__device__ inline f3 volume_unguided_scatter_probability(const VolumeCoefficients &c, float distance,
                                                         bool terminate) {
    return (!terminate && volume_scatters(c)) ? V(1.0f, 1.0f, 1.0f) - volume_transmittance(c.sigma_t, distance)
                                              : V(0.0f, 0.0f, 0.0f);
}

// The scatter-versus-transmit estimator over one segment of `distance` with
// per-channel scatter probability `scatter_probability`.
__device__ inline HomogeneousSample volume_sample_with_probability(const VolumeCoefficients &c, float distance,
                                                                   f3 throughput, float scatter_random,
                                                                   float channel_random, f3 scatter_probability,
                                                                   bool terminate) {
    const f3 one = V(1.0f, 1.0f, 1.0f);
    const f3 segment_t = volume_transmittance(c.sigma_t, distance);
    const f3 segment_emission = throughput * volume_emission_integral(c, distance);
    const bool eligible = !terminate && volume_scatters(c) && distance > 0.0f;
    const f3 sp = eligible ? scatter_probability : V(0.0f, 0.0f, 0.0f);
    const f3 albedo = vol_safe_divide(c.sigma_s, c.sigma_t);
    const f3 ms_albedo = albedo * (one - segment_t) * throughput;
    const VolumeChannelSample ch = volume_sample_channel(ms_albedo + segment_t, throughput, channel_random);
    const float channel_sp = vol_channel(sp, ch.channel);
    const bool scattered = eligible && scatter_random < channel_sp;
    const float rescaled_scatter = scatter_random / (scattered ? channel_sp : 1.0f);
    const float rate = vol_channel(c.sigma_t, ch.channel);
    const float sampled = volume_bounded_exponential_sample(rescaled_scatter, rate, 0.0f, distance);
    const float scatter_pdf = dot(volume_bounded_exponential_pdf(sampled, c.sigma_t, 0.0f, distance) * sp, ch.pdf);
    const f3 scatter_tp =
        throughput * vol_safe_divide(c.sigma_s * volume_transmittance(c.sigma_t, sampled), scatter_pdf);
    const float transmit_pdf = dot(one - sp, ch.pdf);
    const f3 transmit_tp = throughput * vol_safe_divide(segment_t, transmit_pdf);
    const float transmit_rescaled = (scatter_random - channel_sp) / ((!scattered && eligible) ? 1.0f - channel_sp : 1.0f);
    HomogeneousSample r;
    r.transmittance = segment_t;
    r.emission = segment_emission;
    r.throughput = scattered ? scatter_tp : transmit_tp;
    r.scatter_probability = sp;
    r.channel_pdf = ch.pdf;
    r.distance = scattered ? sampled : distance;
    r.event_pdf = scattered ? scatter_pdf : transmit_pdf;
    r.scatter_random = scattered ? rescaled_scatter : transmit_rescaled;
    r.reservoir_random = ch.random;
    r.channel = ch.channel;
    r.scattered = scattered;
    r.active = c.has_extinction || c.has_scatter || c.has_emission;
    return r;
}

__device__ inline HomogeneousSample volume_sample(const VolumeCoefficients &c, float distance, f3 throughput,
                                                  float scatter_random, float channel_random, bool terminate) {
    return volume_sample_with_probability(c, distance, throughput, scatter_random, channel_random,
                                          volume_unguided_scatter_probability(c, distance, terminate), terminate);
}

// Scatter probability of one segment from the primary ray's guide: the
// attenuation alone without guiding, a defensive mix of attenuation and the
// guide's scattered share otherwise.
__device__ inline f3 volume_scatter_probability(const VolumeCoefficients &c, float distance, bool terminate,
                                                const ScatterGuide &g) {
    const bool attenuation_only = terminate || !c.has_scatter || !vol_any_nonzero(c.sigma_s);
    const f3 one = V(1.0f, 1.0f, 1.0f);
    const f3 attenuation = one - vol_exp(vol_neg(c.sigma_t) * distance);
    const f3 radiance = g.transmitted_radiance + g.scattered_radiance;
    const bool history_empty = !vol_any_nonzero(radiance);
    const f3 empty_p = V(c.sigma_t.x > 0.0f ? 0.5f : 0.0f, c.sigma_t.y > 0.0f ? 0.5f : 0.0f,
                         c.sigma_t.z > 0.0f ? 0.5f : 0.0f);
    const float optical = 1.0f - expf(-g.majorant_optical_depth);
    const float safe_optical = optical != 0.0f ? optical : 1.0f;
    const float scale = fmaxf(attenuation.x, fmaxf(attenuation.y, attenuation.z)) / safe_optical;
    const f3 h = vol_safe_divide(g.scattered_radiance, radiance) * scale;
    const f3 history_p = V(fminf(fmaxf(h.x, 0.0f), 1.0f), fminf(fmaxf(h.y, 0.0f), 1.0f), fminf(fmaxf(h.z, 0.0f), 1.0f));
    const f3 guided = history_empty ? empty_p : history_p;
    const f3 defensive = attenuation * 0.25f + guided * 0.75f;
    const f3 probability = g.enabled ? defensive : attenuation;
    return attenuation_only ? V(0.0f, 0.0f, 0.0f) : probability;
}

__device__ inline DirectSamplingState volume_direct_prepare(uint32_t requested, float random, bool enabled) {
    const bool use_mis = requested == kVolumeSampleMis;
    const bool choose_distance = random < 0.5f;
    const uint32_t mis_method = choose_distance ? kVolumeSampleDistance : kVolumeSampleEquiangular;
    const float mis_random = choose_distance ? random * 2.0f : (random - 0.5f) * 2.0f;
    const uint32_t method = use_mis ? mis_method : requested;
    const bool active = enabled && requested != kVolumeSampleNone;
    return DirectSamplingState{active ? method : kVolumeSampleNone, use_mis ? mis_random : random,
                               active && use_mis, active};
}

struct EquiangularGeometry {
    float delta, perpendicular, theta_min, theta_max, theta_range;
    bool valid, uniform;
};

__device__ inline EquiangularGeometry equiangular_geometry(f3 origin, f3 direction, f3 light, float minimum,
                                                           float maximum) {
    const f3 offset = light - origin;
    const float delta = dot(offset, direction);
    const f3 perp = offset - direction * delta;
    const float pd = sqrtf(dot(perp, perp));
    const bool valid = minimum < maximum && pd != 0.0f;
    const float safe = valid ? pd : 1.0f;
    EquiangularGeometry g;
    g.delta = delta;
    g.perpendicular = pd;
    g.theta_min = atan2f(minimum - delta, safe);
    g.theta_max = atan2f(maximum - delta, safe);
    g.theta_range = g.theta_max - g.theta_min;
    g.valid = valid;
    g.uniform = valid && g.theta_range < 1.0e-6f;
    return g;
}

__device__ __forceinline__ float uniform_interval_pdf(float minimum, float maximum) {
    const float length = maximum - minimum;
    return length != 0.0f ? 1.0f / length : 0.0f;
}

__device__ inline EquiangularSample volume_sample_equiangular(f3 origin, f3 direction, f3 light, float minimum,
                                                              float maximum, float random) {
    const EquiangularGeometry g = equiangular_geometry(origin, direction, light, minimum, maximum);
    const float uniform_distance = minimum + random * (maximum - minimum);
    const float angle = random * g.theta_max + (1.0f - random) * g.theta_min;
    const float offset = g.perpendicular * tanf(angle);
    const float sampled = fminf(fmaxf(g.delta + offset, minimum), maximum);
    const float denominator = g.theta_range * (g.perpendicular * g.perpendicular + offset * offset);
    const float analytic = g.perpendicular / (denominator != 0.0f ? denominator : 1.0f);
    const float pdf = g.uniform ? uniform_interval_pdf(minimum, maximum) : analytic;
    return EquiangularSample{g.valid ? (g.uniform ? uniform_distance : sampled) : 0.0f, g.valid ? pdf : 0.0f,
                             g.valid};
}

__device__ inline float volume_equiangular_pdf(f3 origin, f3 direction, f3 light, float minimum, float maximum,
                                               float distance) {
    const EquiangularGeometry g = equiangular_geometry(origin, direction, light, minimum, maximum);
    const float offset = distance - g.delta;
    const float denominator = g.theta_range * (g.perpendicular * g.perpendicular + offset * offset);
    const float analytic = g.perpendicular / (denominator != 0.0f ? denominator : 1.0f);
    const float pdf = g.uniform ? uniform_interval_pdf(minimum, maximum) : analytic;
    return g.valid ? pdf : 0.0f;
}

__device__ __forceinline__ float volume_power_heuristic(float selected, float competing) {
    const float a = selected * selected, b = competing * competing;
    const float sum = a + b;
    return sum > 0.0f ? a / sum : 0.0f;
}

struct HomogeneousDirectSample {
    f3 transmittance, throughput, channel_pdf;
    float distance, distance_pdf, equiangular_pdf, mis_weight;
    uint32_t channel, method;
    bool use_mis, scattered;
};

// The direct-light scatter point on a segment [minimum, maximum] of a ray
// towards a light at `light`: free-flight distance sampling, equiangular
// sampling, or both under the power heuristic, as `sampling` (from
// volume_direct_prepare) selects (HomogeneousVolumeTransport::sample_direct,
// src/luisa/homogeneous_volume_transport.cpp:386-566).  `distance` is the
// segment length whose transmittance weights the channel pick.
__device__ inline HomogeneousDirectSample volume_sample_direct(const VolumeCoefficients &c, float distance,
                                                               f3 throughput, float scatter_random,
                                                               float reservoir_random,
                                                               const DirectSamplingState &sampling, f3 origin,
                                                               f3 direction, f3 light, float minimum,
                                                               float maximum) {
    const f3 one = V(1.0f, 1.0f, 1.0f);
    const f3 segment_t = volume_transmittance(c.sigma_t, distance);
    const bool scattered = sampling.enabled && volume_scatters(c) && distance > 0.0f && minimum < maximum;
    const f3 albedo = vol_safe_divide(c.sigma_s, c.sigma_t);
    // the second throughput factor is part of Cycles' channel measure
    const f3 volume_albedo = albedo * (one - segment_t) * throughput;
    const VolumeChannelSample ch = volume_sample_channel(volume_albedo, throughput, reservoir_random);
    const float rate = vol_channel(c.sigma_t, ch.channel);
    const float distance_sample = volume_bounded_exponential_sample(scatter_random, rate, minimum, maximum);
    const float distance_sample_pdf =
        dot(volume_bounded_exponential_pdf(distance_sample, c.sigma_t, minimum, maximum), ch.pdf);
    const EquiangularSample eq = volume_sample_equiangular(origin, direction, light, minimum, maximum, sampling.random);
    const bool equiangular = sampling.method == kVolumeSampleEquiangular;
    const float sampled = equiangular ? eq.distance : distance_sample;
    const float distance_pdf =
        equiangular ? dot(volume_bounded_exponential_pdf(eq.distance, c.sigma_t, minimum, maximum), ch.pdf)
                    : distance_sample_pdf;
    const float equiangular_pdf =
        equiangular ? eq.pdf : volume_equiangular_pdf(origin, direction, light, minimum, maximum, distance_sample);
    const float selected_pdf = equiangular ? equiangular_pdf : distance_pdf;
    const float competing_pdf = equiangular ? distance_pdf : equiangular_pdf;
    const float mis = sampling.use_mis ? 2.0f * volume_power_heuristic(selected_pdf, competing_pdf) : 1.0f;
    const bool valid = scattered && (!equiangular || eq.valid) && selected_pdf > 0.0f;
    const f3 sampled_t = volume_transmittance(c.sigma_t, sampled);
    const f3 direct_tp = throughput * vol_safe_divide(c.sigma_s * sampled_t, selected_pdf) * mis;
    HomogeneousDirectSample r;
    r.transmittance = valid ? sampled_t : one;
    r.throughput = valid ? direct_tp : V(0.0f, 0.0f, 0.0f);
    r.channel_pdf = ch.pdf;
    r.distance = valid ? sampled : 0.0f;
    r.distance_pdf = valid ? distance_pdf : 0.0f;
    r.equiangular_pdf = valid ? equiangular_pdf : 0.0f;
    r.mis_weight = valid ? mis : 0.0f;
    r.channel = ch.channel;
    r.method = sampling.method;
    r.use_mis = sampling.use_mis;
    r.scattered = valid;
    return r;
}

// ---- phase functions ------------------------------------------------------
// Henyey-Greenstein and Rayleigh scattering at a volume collision
// (include/psycles/luisa/cycles_volume_phase.h: clamp_anisotropy :46-50,
// henyey_greenstein_pdf :121-139, rayleigh_pdf :140-143, sample_direction
// :280-292, sample_henyey_greenstein_cosine :294-312).

__device__ __forceinline__ float phase_clamp_anisotropy(float g) {
    const float s = g > 0.0f ? 1.0f : (g < 0.0f ? -1.0f : 0.0f);
    return s * fminf(fabsf(g), 1.0f - 1.0e-3f);
}

__device__ inline float phase_hg_pdf(float cosine, float g) {
    float result = 0.07957747154594766788f;  // 1 / (4 pi)
    if (fabsf(g) >= 1.0e-3f) {
        const float factor = 1.0f + g * (g - 2.0f * cosine);
        result = (1.0f - g * g) / (4.0f * 3.14159265358979323846f * factor * sqrtf(fmaxf(factor, 0.0f)));
    }
    return result;
}

__device__ __forceinline__ float phase_rayleigh_pdf(float cosine) {
    return (0.1875f * 0.31830988618379067154f) * (1.0f + cosine * cosine);
}

__device__ inline float phase_hg_sample_cosine(float g, float random) {
    float cosine = 1.0f - 2.0f * random;
    if (fabsf(g) >= 1.0e-3f) {
        const float ratio = (1.0f - g * g) / (1.0f - g * cosine);
        cosine = (1.0f + g * g - ratio * ratio) / (2.0f * g);
    }
    return cosine;
}

// 1 / cbrt(value): a bit-pattern first guess and two Newton steps
// (Cycles' fast_inv_cbrtf; include/psycles/luisa/cycles_fast_math.h:229-243)
__device__ __forceinline__ float phase_inverse_cube_root(float value) {
    float r = __int_as_float(0x54a24242 - __float_as_int(value) / 3);
    r = (2.0f / 3.0f) * r + 1.0f / (3.0f * r * r * value);
    r = (2.0f / 3.0f) * r + 1.0f / (3.0f * r * r * value);
    return r;
}

// sample_rayleigh_cosine (cycles_volume_phase.h:313-320): the cubic's root
__device__ __forceinline__ float phase_rayleigh_sample_cosine(float random) {
    const float a = 2.0f - 4.0f * random;
    const float inverse_u = -phase_inverse_cube_root(sqrtf(1.0f + a * a) + a);
    return 1.0f / inverse_u - inverse_u;
}

// the direction at `cosine` from `axis`, azimuth 2 pi * azimuth_sample
__device__ inline f3 phase_direction(f3 axis, float cosine, float azimuth_sample) {
    const float phi = 2.0f * 3.14159265358979323846f * azimuth_sample;
    const float sine = sqrtf(fmaxf(1.0f - cosine * cosine, 0.0f));
    const f3 local = V(sine * cosf(phi), sine * sinf(phi), cosine);
    f3 t, b;
    make_orthonormals(axis, t, b);
    return t * local.x + b * local.y + axis * local.z;
}

// ---- light intervals ------------------------------------------------------
// The part of a ray segment from which a spot or area light can be seen,
// which bounds the equiangular proposal (VolumeLightInterval,
// src/luisa/volume_light_interval.cpp:1-647): a quadratic solve, slab, box,
// plane, elliptic-cylinder and cone clipping, then the spot cone (apex moved
// back for a finite radius, in the lamp's possibly non-uniform frame) and the
// area light's spread cone or zero-spread prism on its emitting side.

struct LightInterval {
    float minimum, maximum;
    bool valid;
};

__device__ __forceinline__ float interval_min(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float interval_max(float a, float b) { return b > a ? b : a; }

struct QuadraticRoots {
    float minimum, maximum;
    bool valid;
};

__device__ inline QuadraticRoots interval_quadratic(float a, float b, float c) {
    const bool valid_linear = a == 0.0f && b != 0.0f;
    const float linear_root = -c / (b != 0.0f ? b : 1.0f);
    const float discriminant = b * b - 4.0f * a * c;
    const bool valid_quadratic = a != 0.0f && discriminant > -1.0e-5f;
    const float root = sqrtf(interval_max(discriminant, 0.0f));
    const float temporary = -0.5f * (b + copysignf(root, b));
    const float r1 = temporary / (a != 0.0f ? a : 1.0f);
    const float r2 = temporary != 0.0f ? c / temporary : r1;
    const float qmin = r2 < r1 ? r2 : r1, qmax = r2 > r1 ? r2 : r1;
    return QuadraticRoots{valid_quadratic ? qmin : linear_root, valid_quadratic ? qmax : linear_root,
                          valid_linear || valid_quadratic};
}

__device__ __forceinline__ LightInterval interval_and(float lo0, float hi0, float lo1, float hi1, bool valid) {
    const float lo = interval_max(lo0, lo1), hi = interval_min(hi0, hi1);
    return LightInterval{lo, hi, valid && lo < hi};
}

__device__ inline LightInterval interval_plane(f3 normal, f3 plane_to_origin, f3 direction, float lo, float hi) {
    const float cosine = dot(direction, normal);
    const float distance = -dot(plane_to_origin, normal) / cosine;
    const float mn = cosine > 0.0f ? interval_max(lo, distance) : lo;
    const float mx = cosine > 0.0f ? hi : interval_min(hi, distance);
    return LightInterval{mn, mx, mn < mx};
}

__device__ inline LightInterval interval_slab(float bound_lo, float bound_hi, float origin, float direction, float lo,
                                              float hi) {
    const bool moving = direction != 0.0f;
    const float safe = moving ? direction : 1.0f;
    const float a = (bound_lo - origin) / safe, b = (bound_hi - origin) / safe;
    const LightInterval bounded = interval_and(lo, hi, interval_min(a, b), interval_max(a, b), moving);
    const bool inside = !moving && origin >= bound_lo && origin <= bound_hi && lo < hi;
    return moving ? bounded : LightInterval{lo, hi, inside};
}

__device__ inline LightInterval interval_box(f3 bound_lo, f3 bound_hi, f3 origin, f3 direction, float lo, float hi) {
    const LightInterval x = interval_slab(bound_lo.x, bound_hi.x, origin.x, direction.x, lo, hi);
    const LightInterval y = interval_slab(bound_lo.y, bound_hi.y, origin.y, direction.y, x.minimum, x.maximum);
    const LightInterval z = interval_slab(bound_lo.z, bound_hi.z, origin.z, direction.z, y.minimum, y.maximum);
    return LightInterval{z.minimum, z.maximum, x.valid && y.valid && z.valid};
}

__device__ inline LightInterval interval_elliptic_cylinder(f3 origin, f3 direction, float radius_u, float radius_v,
                                                           float lo, float hi) {
    const bool radii = radius_u > 0.0f && radius_v > 0.0f;
    const float su = radii ? radius_u : 1.0f, sv = radii ? radius_v : 1.0f;
    float ox = origin.x / su, oy = origin.y / sv;
    const float dx = direction.x / su, dy = direction.y / sv;
    const float a = dx * dx + dy * dy;
    float b = ox * dx + oy * dy;
    const bool movable = a != 0.0f;
    const float middle = -b / (movable ? a : 1.0f);
    ox = ox + dx * middle;
    oy = oy + dy * middle;
    b = ox * dx + oy * dy;
    const float c = (ox * ox + oy * oy) - 1.0f;
    const QuadraticRoots r = interval_quadratic(a, 2.0f * b, c);
    const LightInterval bounded = interval_and(lo, hi, r.minimum + middle, r.maximum + middle, r.valid);
    return LightInterval{bounded.minimum, bounded.maximum, radii && movable && bounded.valid};
}

__device__ inline LightInterval interval_cone(f3 axis, f3 apex_to_origin, f3 direction, float cos2, float lo,
                                              float hi) {
    const bool plane_limit = cos2 < 1.0e-4f;
    const LightInterval plane = interval_plane(axis, apex_to_origin, direction, lo, hi);
    const float length2 = dot(direction, direction);
    const bool valid_direction = length2 > 0.0f;
    const float inverse_length = 1.0f / sqrtf(interval_max(length2, 1.0e-30f));
    const f3 nd = direction * inverse_length;
    const float ad = dot(axis, nd), ao = dot(axis, apex_to_origin);
    const float a = ad * ad - cos2;
    const float b = 2.0f * (ad * ao - cos2 * dot(nd, apex_to_origin));
    const float c = ao * ao - cos2 * dot(apex_to_origin, apex_to_origin);
    const QuadraticRoots r = interval_quadratic(a, b, c);
    const bool min_valid = ao + r.minimum * ad > 0.0f;
    const bool max_valid = ao + r.maximum * ad > 0.0f;
    const bool neither = !max_valid, only_max = max_valid && !min_valid;
    const float hemi_min = only_max ? r.maximum : (neither ? 0.0f : r.minimum);
    const float hemi_max = only_max ? FLT_MAX : (neither ? r.minimum : r.maximum);
    const LightInterval cone = interval_and(lo, hi, hemi_min * inverse_length, hemi_max * inverse_length,
                                            valid_direction && r.valid && (min_valid || max_valid));
    return plane_limit ? plane : cone;
}

__device__ __forceinline__ f3 interval_cross(f3 a, f3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// VolumeLightInterval::spot: the spot cone in the lamp's frame (columns
// axis * scale, inverted through the adjugate), apex moved back by the
// lamp radius so a finite sphere's cone is covered
__device__ inline LightInterval volume_spot_interval(f3 origin, f3 direction, float lo, float hi, f3 center, f3 ax,
                                                     f3 ay, f3 az, f3 scale, float radius, float spot_angle) {
    const f3 cx = ax * scale.x, cy = ay * scale.y, cz = az * scale.z;
    const f3 adj_x = interval_cross(cy, cz);
    const float det = dot(cx, adj_x);
    const bool transform_valid = fabsf(det) > 1.0e-20f;
    const float inv_det = 1.0f / (transform_valid ? det : 1.0f);
    const f3 row_x = adj_x * inv_det, row_y = interval_cross(cz, cx) * inv_det, row_z = interval_cross(cx, cy) * inv_det;
    const float half_angle = 0.5f * spot_angle;
    const float tangent = tanf(half_angle);
    const float transverse = interval_min(scale.x * scale.x, scale.y * scale.y);
    const float axial = scale.z * scale.z;
    const float denominator = tangent * tangent * transverse;
    const bool geometry_valid = transform_valid && denominator > 0.0f;
    const float apex_shift = radius * sqrtf(1.0f + axial / (geometry_valid ? denominator : 1.0f));
    const f3 shifted = origin + vol_neg(az) * apex_shift - center;
    const f3 lo3 = V(dot(row_x, shifted), dot(row_y, shifted), dot(row_z, shifted));
    const f3 ld = V(dot(row_x, direction), dot(row_y, direction), dot(row_z, direction));
    const float ch = cosf(half_angle);
    LightInterval r = interval_cone(V(0.0f, 0.0f, -1.0f), lo3, ld, ch * ch, lo, hi);
    r.valid = r.valid && geometry_valid;
    return r;
}

// VolumeLightInterval::area: the spread cone whose apex sits behind the
// light by its maximal extent / tan(spread / 2), or, for a spread below
// 1e-5, the prism (box or elliptic cylinder) over the light; then the
// emitting side of its plane
__device__ inline LightInterval volume_area_interval(f3 ray_origin, f3 direction, float lo, float hi, f3 center,
                                                     f3 au, f3 av, f3 az, float length_u, float length_v,
                                                     float spread, bool ellipse) {
    const f3 normal = vol_neg(az);
    const f3 origin = ray_origin - center;
    const f3 lorig = V(dot(origin, au), dot(origin, av), dot(origin, normal));
    const f3 ldir = V(dot(direction, au), dot(direction, av), dot(direction, normal));
    const float half_u = 0.5f * length_u, half_v = 0.5f * length_v;
    const float tangent = spread == 3.14159265358979323846f ? FLT_MAX : tanf(0.5f * interval_max(spread, 0.0f));
    const bool nearly_parallel = tangent < 1.0e-5f;
    const LightInterval cylinder = interval_elliptic_cylinder(lorig, ldir, half_u, half_v, lo, hi);
    const LightInterval box = interval_box(V(-half_u, -half_v, 0.0f), V(half_u, half_v, FLT_MAX), lorig, ldir, lo, hi);
    const LightInterval zero_spread = ellipse ? cylinder : box;
    const float extent = 0.5f * (ellipse ? interval_max(length_u, length_v)
                                         : sqrtf(length_u * length_u + length_v * length_v));
    const float safe_tangent = tangent != 0.0f ? tangent : 1.0f;
    const f3 apex_to_origin = origin + normal * (extent / safe_tangent);
    const LightInterval cone = interval_cone(normal, apex_to_origin, direction, 1.0f / (1.0f + tangent * tangent), lo, hi);
    const LightInterval region = nearly_parallel ? zero_spread : cone;
    const LightInterval side = interval_plane(normal, origin, direction, region.minimum, region.maximum);
    return LightInterval{side.minimum, side.maximum, region.valid && side.valid};
}

}  // namespace psx
