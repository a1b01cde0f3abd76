// psx_host.cpp - host half of the MI355X render path and its C ABI.
//
// Implements include/psx.h: scene compilation (the flat SceneSnapshot is
// lowered to device tables and a two-level BVH4), render sessions with film
// buffers resident in HBM, the ordered sample-batch dispatch loop and the
// write_passes normalisation.  Reference counterparts are cited per function.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <array>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/psx.h"
#include "psx_bvh.h"
#include "psx_device.h"
#include "psx_light_tree.h"
#include "psx_rtc.h"

namespace psx {

// psx_kernels.hip
uint32_t megakernel_lanes();
hipError_t launch_render(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                         const FilmView &film, uint32_t first, uint32_t count, Counters *counters, bool count_nodes,
                         uint32_t *spill, uint32_t *cursor, hipStream_t stream);
hipError_t launch_trace(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                        uint32_t sample, uint32_t *records, uint32_t *spill, hipStream_t stream);
hipError_t launch_path_trace(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                             uint32_t sample, uint32_t px, uint32_t py, float4 *slots, uint32_t *records,
                             uint32_t *spill, hipStream_t stream);
hipError_t launch_primary(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                          uint32_t sample, uint32_t *records, uint32_t *spill, hipStream_t stream);
hipError_t launch_sampler(const KernelParams &k, const float4 *sobol, uint32_t sample, uint32_t first_dim,
                          uint32_t dim_count, uint32_t *out, hipStream_t stream);
hipError_t launch_material_tables(const SceneView &s, float4 *shadow, float4 *emission, uint32_t count,
                                  hipStream_t stream);
hipError_t launch_instance_opacity(const SceneView &s, uint32_t *flags, uint32_t count, hipStream_t stream);
// psx_query.hip
uint32_t query_resident_lanes();
uint32_t query_lanes(uint32_t count, uint32_t max_lanes);
hipError_t launch_intersect(const SceneView &s, const psx_ray_query *queries, uint32_t count, psx_ray_hit *hits,
                            uint32_t *cursor, uint32_t *spill, uint32_t lanes, hipStream_t stream);
hipError_t launch_probe(uint32_t function, const float *inputs, uint32_t in_width, uint32_t count, float *outputs,
                        uint32_t out_width, hipStream_t stream);
hipError_t setup_noalloc_scratch();
hipError_t launch_intersect_noalloc(const SceneView &s, const psx_ray_query *queries, uint32_t count,
                                    psx_ray_hit *hits, hipStream_t stream);
// psx_wavefront.hip
uint32_t wavefront_lanes();
uint32_t wavefront_padded_pixels(const KernelParams &k, uint32_t tile_shift);
uint32_t wavefront_queue_headroom();
uint32_t wavefront_list_headroom();
hipError_t launch_wavefront(const SceneView &s, const KernelParams &k, const float4 *sobol, const float *filter,
                            const FilmView &film, const WavefrontView &w, uint32_t first, uint32_t count,
                            const uint32_t *first_at, Counters *counters, bool count_nodes, uint32_t *spill,
                            hipEvent_t *events, const GeneratedShade *generated, hipStream_t stream);
hipError_t launch_store_word(uint32_t *dst, uint32_t value, hipStream_t stream);
uint32_t shade_variant(const SceneView &s, bool light_passes);

namespace {

struct DeviceError {
    int status;
};

// the last HIP failure on this thread, as "<error name> at psx_host.cpp:<line>"
thread_local std::string g_last_error;

void hip_check(hipError_t e, int line = __builtin_LINE()) {
    if (e == hipSuccess) return;
    g_last_error = std::string(hipGetErrorName(e)) + " at psx_host.cpp:" + std::to_string(line);
    if (e == hipErrorOutOfMemory) throw DeviceError{PSX_ERROR_OUT_OF_MEMORY};
    throw DeviceError{PSX_ERROR_DEVICE};
}

// RAII device allocation
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { reset(); }
    void reset() {
        if (ptr_) {
            if (stream_)
                (void)hipFreeAsync(ptr_, stream_);
            else
                (void)hipFree(ptr_);
        }
        ptr_ = nullptr;
        bytes_ = 0;
    }
    void allocate(size_t bytes) {
        reset();
        if (bytes == 0) bytes = 16;
        if (stream_)
            hip_check(hipMallocAsync(&ptr_, bytes, stream_));
        else
            hip_check(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    // Stream-ordered allocation and release on `stream` from now on: no
    // device-wide synchronisation when the buffer is freed while other
    // streams run.  The stream must outlive the buffer's allocation.
    void set_stream(hipStream_t stream) {
        reset();
        stream_ = stream;
    }
    template <typename T>
    void upload(const std::vector<T> &v) {
        allocate(v.size() * sizeof(T));
        if (v.empty()) return;
        if (stream_) {
            hip_check(hipMemcpyAsync(ptr_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
            hip_check(hipStreamSynchronize(stream_));
        } else {
            hip_check(hipMemcpy(ptr_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        }
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(ptr_);
    }
    size_t bytes() const { return bytes_; }

  private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
    hipStream_t stream_ = nullptr;
};

// ---------------------------------------------------------------------------
// host math shared by scene compilation (same contracts as the kernels)
// ---------------------------------------------------------------------------
struct Vec3 {
    float x, y, z;
};
inline Vec3 v3(float x, float y, float z) { return Vec3{x, y, z}; }
inline Vec3 sub(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 cross(Vec3 a, Vec3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 divs(Vec3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }

// transform_point (src/luisa/path_tracer_scene_geometry.cpp:93-113)
inline Vec3 transform_point(const float *e, Vec3 p) {
    return v3(std::fma(p.x, e[0], std::fma(p.y, e[4], std::fma(p.z, e[8], e[12]))),
              std::fma(p.x, e[1], std::fma(p.y, e[5], std::fma(p.z, e[9], e[13]))),
              std::fma(p.x, e[2], std::fma(p.y, e[6], std::fma(p.z, e[10], e[14]))));
}

// cycles_inverse_transform (src/luisa/path_tracer_scene_geometry.cpp:231-268)
void cycles_inverse(const float *e, float *r) {
    Vec3 x = v3(e[0], e[1], e[2]), y = v3(e[4], e[5], e[6]), z = v3(e[8], e[9], e[10]);
    const Vec3 t = v3(e[12], e[13], e[14]);
    float det = dot(x, cross(y, z));
    if (det == 0.0f) {
        x.x += 1.0e-8f;
        y.y += 1.0e-8f;
        z.z += 1.0e-8f;
        det = dot(x, cross(y, z));
        if (det == 0.0f) det = std::numeric_limits<float>::max();
    }
    const Vec3 ix = divs(cross(y, z), det), iy = divs(cross(z, x), det), iz = divs(cross(x, y), det);
    const float vals[16] = {ix.x, iy.x, iz.x, 0.0f, ix.y, iy.y, iz.y, 0.0f, ix.z, iy.z, iz.z, 0.0f,
                            -dot(ix, t), -dot(iy, t), -dot(iz, t), 1.0f};
    std::memcpy(r, vals, sizeof(vals));
}

Affine to_affine(const float *e) {
    Affine a;
    const int src[12] = {0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14};
    for (int i = 0; i < 12; ++i) a.m[i] = e[src[i]];
    return a;
}

void transpose4(const float *a, float *r) {
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) r[c * 4 + row] = a[row * 4 + c];
}

// ---------------------------------------------------------------------------
// tabulated Sobol LUT (src/sampling/tabulated_sobol.cpp:24-199)
// ---------------------------------------------------------------------------
const uint32_t kSobolXors[4][32] = {
    {0u},
    {0x00000000u, 0x00000001u, 0x00000001u, 0x00000007u, 0x00000001u, 0x00000013u, 0x00000015u, 0x0000007fu,
     0x00000001u, 0x00000103u, 0x00000105u, 0x0000070fu, 0x00000111u, 0x00001333u, 0x00001555u, 0x00007fffu,
     0x00000001u, 0x00010003u, 0x00010005u, 0x0007000fu, 0x00010011u, 0x00130033u, 0x00150055u, 0x007f00ffu,
     0x00010101u, 0x01030303u, 0x01050505u, 0x070f0f0fu, 0x01111111u, 0x13333333u, 0x15555555u, 0x7fffffffu},
    {0x00000000u, 0x00000001u, 0x00000003u, 0x00000001u, 0x00000005u, 0x0000001fu, 0x0000002bu, 0x0000003du,
     0x00000011u, 0x00000133u, 0x00000377u, 0x00000199u, 0x00000445u, 0x00001ccfu, 0x00002ddbu, 0x0000366du,
     0x00000101u, 0x00010303u, 0x00030707u, 0x00010909u, 0x00051515u, 0x001f3f3fu, 0x002b6b6bu, 0x003dbdbdu,
     0x00101011u, 0x01303033u, 0x03707077u, 0x01909099u, 0x04515145u, 0x1cf3f3cfu, 0x2db6b6dbu, 0x36dbdb6du},
    {0x00000000u, 0x00000001u, 0x00000000u, 0x00000003u, 0x0000000du, 0x0000000cu, 0x00000005u, 0x0000004fu,
     0x00000014u, 0x000000e7u, 0x00000329u, 0x0000039cu, 0x00000011u, 0x00001033u, 0x00000044u, 0x000030bbu,
     0x0000d1cdu, 0x0000c2ecu, 0x00005415u, 0x0004fc3fu, 0x00015054u, 0x000e5c97u, 0x0032e5b9u, 0x0039725cu,
     0x00000101u, 0x01000303u, 0x00000404u, 0x03000b0bu, 0x0d001d1du, 0x0c002c2cu, 0x05004545u, 0x4f00cfcfu}};

uint32_t hash_hp(uint32_t v) {
    v ^= v >> 16u;
    v *= 0x21f0aaadu;
    v ^= v >> 15u;
    v *= 0xd35a2d97u;
    v ^= v >> 15u;
    return v ^ 0xe6fe3bebu;
}

float hash_hp_float(uint32_t v) { return static_cast<float>(hash_hp(v)) * (1.0f / 4294967808.0f); }

uint32_t sequence_size_for(uint32_t samples) {
    const uint32_t c = std::min(std::max(samples, 1u), 8192u);
    uint32_t u = 1u;
    if (c != 1u) u = 1u << (32u - static_cast<uint32_t>(__builtin_clz(c - 1u)));
    return std::min(std::max(u, 256u), 8192u);
}

std::vector<float> sobol_table(uint32_t size) {
    std::vector<float> t(static_cast<size_t>(size) * 256u * 4u);
    for (uint32_t pattern = 0; pattern < 256u; ++pattern) {
        float *p = t.data() + static_cast<size_t>(pattern) * size * 4u;
        // hash_hp_seeded_uint(pattern, 0x44605a73): the seed is pre-mixed
        uint32_t seed = 0x44605a73u;
        seed ^= seed << 19u;
        uint32_t rng = hash_hp(pattern ^ seed);
        for (int c = 0; c < 4; ++c) p[c] = hash_hp_float(rng++);
        uint32_t log_n = 0;
        for (uint32_t n = 1; n < size; ++log_n, n *= 2u) {
            const float strata = static_cast<float>(n * 2u);
            for (uint32_t i = 0; i < n && n + i < size; ++i) {
                uint32_t occ[4];
                for (int c = 0; c < 4; ++c)
                    occ[c] = static_cast<uint32_t>(p[4u * (i ^ kSobolXors[c][log_n]) + c] * strata);
                for (int c = 0; c < 4; ++c)
                    p[4u * (n + i) + c] = (static_cast<float>(occ[c] ^ 1u) + hash_hp_float(rng++)) / strata;
            }
        }
    }
    return t;
}

// ---------------------------------------------------------------------------
// pixel filter LUT (src/sampling/pixel_filter.cpp:30-86)
// ---------------------------------------------------------------------------
std::vector<float> pixel_filter_table(uint32_t filter, float width) {
    constexpr size_t size = 1024;
    width = std::max(width, 1.0e-5f);
    if (filter == PSX_FILTER_GAUSSIAN) width *= 3.0f;
    if (filter == PSX_FILTER_BLACKMAN_HARRIS) width *= 2.0f;
    auto eval = [&](float value) -> float {
        if (filter == PSX_FILTER_GAUSSIAN) {
            value *= 6.0f / width;
            return std::exp(-2.0f * value * value);
        }
        if (filter == PSX_FILTER_BLACKMAN_HARRIS) {
            constexpr float two_pi = 6.28318530717958647692f;
            value = two_pi * (value / width + 0.5f);
            return 0.35875f - 0.48829f * std::cos(value) + 0.14128f * std::cos(2.0f * value) -
                   0.01168f * std::cos(3.0f * value);
        }
        return 1.0f;
    };
    const size_t resolution = size - 1u;
    std::vector<float> cdf(size, 0.0f);
    const float range = width * 0.5f;
    for (size_t i = 0; i < resolution; ++i) {
        const float value = range * static_cast<float>(i) / static_cast<float>(resolution - 1u);
        cdf[i + 1u] = cdf[i] + std::fabs(eval(value));
    }
    const float norm = cdf[resolution] == 0.0f ? 0.0f : 1.0f / cdf[resolution];
    for (auto &v : cdf) v *= norm;
    cdf[resolution] = 1.0f;
    std::vector<float> inv(size, 0.0f);
    const size_t half = (size - 1u) / 2u;
    for (size_t i = 0; i <= half; ++i) {
        const float x = static_cast<float>(i) / static_cast<float>(half);
        size_t index = static_cast<size_t>(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
        float t = 0.0f;
        if (index < resolution) {
            t = (x - cdf[index]) / (cdf[index + 1u] - cdf[index]);
        } else {
            index = resolution;
        }
        const float y = (static_cast<float>(index) + t) / static_cast<float>(size - 1u) * (2.0f * range);
        inv[half + i] = 0.5f * (1.0f + y);
        inv[half - i] = 0.5f * (1.0f - y);
    }
    return inv;
}

// ---------------------------------------------------------------------------
// light distribution (src/sampling/light_distribution.cpp:25-138)
// ---------------------------------------------------------------------------
// Emissive triangles weighted by world area, then `lamps` analytic lamps;
// when both families exist each takes half of the probability.
uint32_t build_distribution(const std::vector<float> &areas, uint32_t lamps, std::vector<DistEntry> &e,
                            float &tri_pdf, float &light_pdf) {
    auto valid = [](float a) { return std::isfinite(a) && a > 0.0f ? a : 0.0f; };
    const uint32_t n = static_cast<uint32_t>(areas.size());
    const uint32_t count = n + lamps;
    e.assign(count + 1u, DistEntry{});
    float total = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        e[i] = DistEntry{total, 0.0f, 0u, i};
        total += valid(areas[i]);
    }
    const float triangle_area = total;
    const float lamp_area = lamps > 0u ? (total > 0.0f ? total / static_cast<float>(lamps) : 1.0f) : 0.0f;
    for (uint32_t j = 0; j < lamps; ++j) {
        e[n + j] = DistEntry{total, 0.0f, 1u, j};
        total += lamp_area;
    }
    e[count] = DistEntry{total, 0.0f, 3u, 0u};
    float tp = 0.0f, lp = 0.0f;
    if (triangle_area > 0.0f) {
        tp = 1.0f / triangle_area;
        if (lamps > 0u) tp *= 0.5f;
    }
    if (lamps > 0u) {
        lp = 1.0f / static_cast<float>(lamps);
        if (triangle_area > 0.0f) lp *= 0.5f;
    }
    for (uint32_t i = 0; i < n; ++i) e[i].selection_pdf = valid(areas[i]) * tp;
    for (uint32_t i = n; i < count; ++i) e[i].selection_pdf = lp;
    if (total > 0.0f) {
        const float inv = 1.0f / total;
        for (uint32_t i = 0; i < count; ++i) e[i].cumulative *= inv;
        e[count].cumulative = 1.0f;
    }
    tri_pdf = tp;
    light_pdf = lp;
    return total > 0.0f ? count : 0u;
}

// ---------------------------------------------------------------------------
// materials: SurfaceProgram -> flat leaf list
// (compile_surface_program / bind_surface_parameters,
//  include/psycles/compiler/surface_program.h:635-647)
// ---------------------------------------------------------------------------
struct HostMaterial {
    std::vector<psx_closure_instr> instr;
    uint32_t root = 0;
    std::vector<float> params;
    uint32_t emission_sampling = 0;
    bool may_emit = false;
    uint64_t signature = 0;
};

bool flatten_program(const HostMaterial &m, uint32_t id, Leaf path, std::vector<Leaf> &out, int guard,
                     std::string &error) {
    if (id >= m.instr.size() || guard > 64) {
        error = "closure instruction index out of range";
        return false;
    }
    const psx_closure_instr &ins = m.instr[id];
    auto param_ok = [&](uint32_t p, uint32_t n) { return p + n <= m.params.size(); };
    switch (ins.op) {
    case PSX_CLOSURE_NULL:
        return true;
    case PSX_CLOSURE_ADD:
        return flatten_program(m, ins.a, path, out, guard + 1, error) &&
               flatten_program(m, ins.b, path, out, guard + 1, error);
    case PSX_CLOSURE_MIX: {
        if (path.depth >= kMaxLeafDepth) {
            error = "mix nesting deeper than supported";
            return false;
        }
        if (!param_ok(ins.factor, 1)) {
            error = "mix factor parameter out of range";
            return false;
        }
        Leaf a = path, b = path;
        a.factor[a.depth] = ins.factor;
        a.depth += 1;
        b.factor[b.depth] = ins.factor;
        b.second_mask |= 1u << b.depth;
        b.depth += 1;
        return flatten_program(m, ins.a, a, out, guard + 1, error) &&
               flatten_program(m, ins.b, b, out, guard + 1, error);
    }
    case PSX_CLOSURE_DIFFUSE:
    case PSX_CLOSURE_EMISSION:
    case PSX_CLOSURE_TRANSPARENT: {
        Leaf l = path;
        l.op = ins.op;
        l.color = ins.color;
        l.aux = ins.op == PSX_CLOSURE_EMISSION ? ins.strength : (ins.op == PSX_CLOSURE_DIFFUSE ? ins.roughness : 0u);
        if (!param_ok(l.color, 3) || !param_ok(l.aux, 1)) {
            error = "closure parameter out of range";
            return false;
        }
        out.push_back(l);
        return true;
    }
    default:
        error = "unsupported closure operation";
        return false;
    }
}

// The material's closure tree as flat leaves, within the supported size.
bool lower_leaves(const HostMaterial &m, std::vector<Leaf> &leaves, std::string &error) {
    leaves.clear();
    if (m.instr.empty()) return true;
    Leaf root{};
    if (!flatten_program(m, m.root, root, leaves, 0, error)) return false;
    uint32_t scattering = 0;
    for (const auto &l : leaves) scattering += l.op != kLeafEmission ? 1u : 0u;
    if (leaves.size() > kMaxLeaves || scattering > 4u) {
        error = "closure tree larger than supported";
        return false;
    }
    return true;
}

HostMaterial host_material(const psx_material &src) {
    HostMaterial m;
    m.instr.assign(src.instructions, src.instructions + src.instruction_count);
    m.root = src.root;
    m.params.assign(src.parameters, src.parameters + src.parameter_count);
    m.emission_sampling = src.emission_sampling;
    return m;
}

// estimate_surface_emission (src/compiler/surface_program.cpp:244-330)
void emission_estimate(const HostMaterial &m, uint32_t id, float out[3], int guard) {
    out[0] = out[1] = out[2] = 0.0f;
    if (id >= m.instr.size() || guard > 64) return;
    const psx_closure_instr &ins = m.instr[id];
    if (ins.op == PSX_CLOSURE_EMISSION) {
        for (int c = 0; c < 3; ++c) out[c] = m.params[ins.color + c] * m.params[ins.strength];
    } else if (ins.op == PSX_CLOSURE_ADD || ins.op == PSX_CLOSURE_MIX) {
        float a[3], b[3];
        emission_estimate(m, ins.a, a, guard + 1);
        emission_estimate(m, ins.b, b, guard + 1);
        const float f = ins.op == PSX_CLOSURE_MIX ? m.params[ins.factor] : 0.0f;
        for (int c = 0; c < 3; ++c)
            out[c] = ins.op == PSX_CLOSURE_MIX ? a[c] * (1.0f - f) + b[c] * f : a[c] + b[c];
    }
}

void refresh_may_emit(HostMaterial &m) {
    float e[3] = {0, 0, 0};
    if (!m.instr.empty()) emission_estimate(m, m.root, e, 0);
    m.may_emit = (e[0] != 0.0f || e[1] != 0.0f || e[2] != 0.0f) && m.emission_sampling != PSX_EMISSION_SAMPLING_NONE;
}

// Binary BVH over lamp bounding spheres for the forward-hit search (nodes in
// depth-first order with skip indices; LampNode).  Leaves hold at most two
// lamps; inner nodes split their lamps at the median centre along the widest
// centre extent.  Boxes are padded well beyond float rounding of the sphere
// bounds and of the device's slab test, so a lamp the exact hit test accepts
// is never culled.
void build_lamp_bvh(const std::vector<float4> &spheres, std::vector<LampNode> &nodes, std::vector<uint32_t> &order) {
    constexpr uint32_t kLeaf = 2u;
    order.resize(spheres.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    nodes.clear();
    std::vector<uint32_t> right;  // right child of each inner node (0 for leaves)
    auto centre = [&](uint32_t i, int a) { return a == 0 ? spheres[i].x : a == 1 ? spheres[i].y : spheres[i].z; };
    std::function<uint32_t(uint32_t, uint32_t)> build = [&](uint32_t first, uint32_t count) -> uint32_t {
        LampNode n{};
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t k = first; k < first + count; ++k) {
            const float4 sp = spheres[order[k]];
            const float r = std::max(sp.w, 0.0f);
            const float pad = 0.05f * r + 1.0e-3f * (std::fabs(sp.x) + std::fabs(sp.y) + std::fabs(sp.z) + r) + 1.0e-6f;
            for (int a = 0; a < 3; ++a) {
                const float c = centre(order[k], a);
                lo[a] = std::min(lo[a], c - r - pad);
                hi[a] = std::max(hi[a], c + r + pad);
                clo[a] = std::min(clo[a], c);
                chi[a] = std::max(chi[a], c);
            }
        }
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = lo[a];
            n.hi[a] = hi[a];
        }
        const uint32_t index = static_cast<uint32_t>(nodes.size());
        nodes.push_back(n);
        right.push_back(0u);
        if (count <= kLeaf) {
            nodes[index].leaf = first | (count << 24);
            return index;
        }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        const uint32_t half = count / 2u;
        std::nth_element(order.begin() + first, order.begin() + first + half, order.begin() + first + count,
                         [&](uint32_t x, uint32_t y) {
                             const float cx = centre(x, axis), cy = centre(y, axis);
                             return cx < cy || (cx == cy && x < y);
                         });
        (void)build(first, half);  // the left child is index + 1
        right[index] = build(first + half, count - half);
        return index;
    };
    if (!order.empty()) (void)build(0u, static_cast<uint32_t>(order.size()));
    std::function<void(uint32_t, uint32_t)> link = [&](uint32_t i, uint32_t miss) {
        nodes[i].miss = miss;
        if (nodes[i].leaf != 0u) return;
        link(i + 1u, right[i]);
        link(right[i], miss);
    };
    if (!nodes.empty()) link(0u, static_cast<uint32_t>(nodes.size()));
}

// Shade sort key of an instance: its materials' common key, or the last key
// when they differ in structure (or it has none).
uint32_t instance_shade_class(const psx_instance &src, const Mesh &mesh, const std::vector<uint32_t> &slots,
                              const std::vector<uint32_t> &material_class) {
    uint32_t key = 0u;
    auto add = [&](uint32_t material) {
        const uint32_t c = material < material_class.size() ? material_class[material] : kShadeKeys - 1u;
        key = key == 0u || key == c ? c : kShadeKeys - 1u;
    };
    for (uint32_t o = 0; o < src.material_override_count; ++o) add(src.material_overrides[o]);
    for (uint32_t k = 0; k < mesh.slot_count; ++k) add(slots[mesh.slot_offset + k]);
    return key == 0u ? kShadeKeys - 1u : key;
}

uint64_t structure_signature(const HostMaterial &m) {
    uint64_t h = 14695981039346656037ull;
    auto mix = [&](uint32_t v) {
        for (int b = 0; b < 4; ++b) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    };
    mix(m.root);
    mix(static_cast<uint32_t>(m.params.size()));
    for (const auto &i : m.instr) {
        mix(i.op);
        mix(i.a);
        mix(i.b);
        mix(i.factor);
        mix(i.color);
        mix(i.strength);
        mix(i.roughness);
    }
    return h;
}

}  // namespace

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct Backend {
    int device = 0;
    psx_backend_options options{};
    ProgramCache programs;  // generated shade kernels (PSX_BACKEND_GENERATED_PROGRAMS)
    // Held while a session captures a dispatch graph, and by the render
    // server's worker while it makes device calls: a synchronous upload
    // (null stream) or module load on another thread during a capture
    // invalidates it, since session streams synchronise with the null
    // stream.
    std::mutex capture_gate;
};

// Light-sampling tables that follow material emission: emitter list, flat
// CDF, light tree.  Shared by parameter versions whose edits left them valid.
struct EmitterTables {
    DeviceBuffer emitters, distribution, tree_nodes, tree_emitters, tree_mapping, emitter_cprim;
    size_t bytes() const {
        return emitters.bytes() + distribution.bytes() + tree_nodes.bytes() + tree_emitters.bytes() +
               tree_mapping.bytes() + emitter_cprim.bytes();
    }
    explicit EmitterTables(hipStream_t stream) {
        for (DeviceBuffer *b : {&emitters, &distribution, &tree_nodes, &tree_emitters, &tree_mapping, &emitter_cprim})
            b->set_stream(stream);
    }
};

// Everything on the device that depends on material parameter blocks.  A
// parameter-only edit builds a new version and leaves the old one intact for
// work still using it; the last holder frees it.
struct ParamTables {
    DeviceBuffer params, materials_buf, material_shadow, material_emission, instance_opaque;
    std::shared_ptr<EmitterTables> light;
    // the version's own tables; `light` may be shared with other versions
    size_t bytes() const {
        return params.bytes() + materials_buf.bytes() + material_shadow.bytes() + material_emission.bytes() +
               instance_opaque.bytes();
    }
    hipEvent_t ready = nullptr;  // recorded after the last upload and table kernel
    explicit ParamTables(hipStream_t stream) {
        for (DeviceBuffer *b : {&params, &materials_buf, &material_shadow, &material_emission, &instance_opaque})
            b->set_stream(stream);
        hip_check(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    }
    ~ParamTables() {
        if (ready) (void)hipEventDestroy(ready);
    }
};

struct Scene {
    Backend *backend = nullptr;
    // host copies needed after compile
    std::vector<HostMaterial> materials;
    std::vector<Material> gpu_materials;
    std::vector<LeafLayout> programs;  // closure-leaf layout of each generated surface program
    std::vector<psx_instance> instances;
    std::vector<std::vector<uint32_t>> overrides;
    std::vector<Mesh> meshes;
    std::vector<std::vector<float>> mesh_positions;
    std::vector<std::vector<uint32_t>> mesh_tris;
    std::vector<std::vector<uint32_t>> mesh_tri_slots;
    std::vector<std::vector<uint32_t>> mesh_slots;
    psx_camera camera{};
    uint32_t max_stack = 0;
    uint32_t tlas_depth = 0, blas_depth = 0;
    // device buffers
    std::vector<Lamp> lamp_list;
    std::vector<uint32_t> mesh_prim_offsets;
    DeviceBuffer nodes, tris, instances_buf, meshes_buf, positions, indices, normals, smooth, slots, tri_slots,
        overrides_buf, leaves, lamps, lamp_spheres, lamp_bits, lamp_nodes, lamp_order;
    // parameter versions are built, and released, on this stream (see
    // DeviceBuffer::set_stream)
    hipStream_t table_stream = nullptr;
    hipStream_t query_stream = nullptr;   // psx_scene_intersect* work
    std::shared_ptr<ParamTables> tables;  // current parameter version
    SceneView view{};                     // geometry + the current version

    ~Scene() {
        tables.reset();
        if (table_stream) {
            (void)hipStreamSynchronize(table_stream);
            (void)hipStreamDestroy(table_stream);
        }
        if (query_stream) {
            (void)hipStreamSynchronize(query_stream);
            (void)hipStreamDestroy(query_stream);
        }
    }
    uint64_t triangles = 0;

    // device bytes of the geometry and lamp tables (parameter versions apart)
    size_t geometry_bytes() const {
        size_t n = 0;
        for (const DeviceBuffer *b : {&nodes, &tris, &instances_buf, &meshes_buf, &positions, &indices, &normals,
                                      &smooth, &slots, &tri_slots, &overrides_buf, &leaves, &lamps, &lamp_spheres,
                                      &lamp_bits, &lamp_nodes, &lamp_order})
            n += b->bytes();
        return n;
    }

    void upload_materials(ParamTables &t);
    // a new parameter version from the host materials; the emitter tables
    // are rebuilt when `emission_changed`, shared otherwise
    void apply_parameter_version(bool emission_changed) {
        auto fresh = std::make_shared<ParamTables>(table_stream);
        upload_materials(*fresh);
        if (emission_changed) {
            rebuild_emitters(*fresh);
        } else {
            fresh->light = tables->light;
        }
        refresh_material_tables(*fresh);
        tables = std::move(fresh);
    }
    void rebuild_emitters(ParamTables &t);

    // Triangles whose material can emit for some parameter block (its
    // program has an emission closure and its emission sampling is not
    // NONE), on light-visible instances, in (instance, primitive) order: the
    // order in which the emitter list enumerates emissive triangles.  Built
    // once, so parameter edits do not rescan every instanced triangle.
    struct EmissiveCandidate {
        uint32_t instance, prim, material;
    };
    std::vector<EmissiveCandidate> emissive_candidates;
    bool candidates_indexed = false;
    void index_emissive_candidates();
    void refresh_material_tables(ParamTables &t);
};

// LightDesc -> LightGpu for point, spot and distant lamps
// (src/luisa/path_tracer_scene.cpp:1735-1870, no lamp shaders)
Lamp make_lamp(const psx_light &l) {
    Lamp g{};
    g.type = l.type;
    float axes[3][3], lengths[3];
    for (int c = 0; c < 3; ++c) {
        const float x = l.transform[4 * c], y = l.transform[4 * c + 1], z = l.transform[4 * c + 2];
        const float n = std::sqrt(x * x + y * y + z * z);
        if (n <= 1.0e-20f) {
            axes[c][0] = axes[c][1] = axes[c][2] = 0.0f;
            lengths[c] = 0.0f;
        } else {
            axes[c][0] = x / n;
            axes[c][1] = y / n;
            axes[c][2] = z / n;
            lengths[c] = n;
        }
    }
    if (axes[2][0] == 0.0f && axes[2][1] == 0.0f && axes[2][2] == 0.0f) axes[2][2] = 1.0f;
    for (int i = 0; i < 3; ++i) {
        g.position[i] = l.transform[12 + i];
        g.axis_x[i] = axes[0][i];
        g.axis_y[i] = axes[1][i];
        g.axis_z[i] = axes[2][i];
        g.axis_scale[i] = lengths[i];
        g.color[i] = l.color[i];
    }
    g.power = l.power;
    g.radius = l.size;
    g.spot_angle = l.spot_angle;
    g.spot_smooth = l.spot_smooth;
    g.angle = l.angle;
    g.visibility = l.visibility_mask;
    g.max_bounces = l.max_bounces;
    g.size_u = l.size * lengths[0];
    g.size_v = (l.size_y > 0.0f ? l.size_y : l.size) * lengths[1];
    g.spread = l.spread;
    // the MIS property only holds where a forward technique exists
    const bool use_mis = (l.flags & PSX_LIGHT_FLAG_USE_MIS) != 0u;
    bool effective_mis;
    if (l.type == PSX_LIGHT_DISTANT)
        effective_mis = use_mis && l.angle > 0.0f;
    else if (l.type == PSX_LIGHT_AREA)
        effective_mis = use_mis && g.size_u != 0.0f && g.size_v != 0.0f && l.spread > 0.0f;
    else
        effective_mis = use_mis && l.size > 0.0f;
    constexpr float kPi = 3.14159265358979f;
    g.flags = ((l.flags & PSX_LIGHT_FLAG_NORMALIZE) ? kLampNormalize : 0u) |
              ((l.flags & PSX_LIGHT_FLAG_ELLIPSE) ? kLampEllipse : 0u) |
              ((l.flags & PSX_LIGHT_FLAG_SPHERE) ? kLampSphere : 0u) | (effective_mis ? kLampUseMis : 0u) |
              (effective_mis && l.type != PSX_LIGHT_DISTANT ? kLampForward : 0u) |
              (l.type == PSX_LIGHT_AREA && l.spread >= kPi - 1.0e-6f ? kLampFullSpread : 0u) |
              ((l.flags & PSX_LIGHT_FLAG_CAST_SHADOW) ? kLampCastShadow : 0u);
    return g;
}

// per-material shadow extinction and constant emission, evaluated on the
// device by the same closure code the kernels run
void Scene::refresh_material_tables(ParamTables &t) {
    const uint32_t n = static_cast<uint32_t>(gpu_materials.size());
    if (t.material_shadow.bytes() < std::max<size_t>(n * sizeof(float4), 16u)) {
        t.material_shadow.allocate(n * sizeof(float4));
        t.material_emission.allocate(n * sizeof(float4));
    }
    view.material_shadow = t.material_shadow.as<float4>();
    view.material_emission = t.material_emission.as<float4>();
    hip_check(launch_material_tables(view, t.material_shadow.as<float4>(), t.material_emission.as<float4>(), n,
                                     table_stream));
    // the shadow-ray early end of each instance follows its materials
    const uint32_t ni = static_cast<uint32_t>(instances.size());
    if (t.instance_opaque.bytes() < std::max<size_t>(ni * sizeof(uint32_t), 16u))
        t.instance_opaque.allocate(ni * sizeof(uint32_t));
    view.instance_opaque = t.instance_opaque.as<uint32_t>();
    hip_check(launch_instance_opacity(view, t.instance_opaque.as<uint32_t>(), ni, table_stream));
    // work of this version waits on `ready` (hipStreamWaitEvent), callers
    // that need the tables on the host side of the call synchronise on it
    hip_check(hipEventRecord(t.ready, table_stream));
}

// parameter blocks (concatenated in material order, at the offsets fixed by
// compilation) and the material records
void Scene::upload_materials(ParamTables &t) {
    std::vector<float> params;
    for (const HostMaterial &m : materials) params.insert(params.end(), m.params.begin(), m.params.end());
    t.params.upload(params);
    t.materials_buf.upload(gpu_materials);
    view.params = t.params.as<float>();
    view.materials = t.materials_buf.as<Material>();
}

// light-tree measures of an emissive triangle and of a lamp
// (make_triangle_light_tree_emitter / make_analytic_light_tree_emitter,
//  src/luisa/path_tracer_light_tree_scene.cpp:130-250)
constexpr float kTreePi = 3.14159265358979323846f;
constexpr float kTreeHalfPi = 0.5f * kTreePi;

Vec3 tree_unit_or(Vec3 v, Vec3 fallback) {
    const float n = std::sqrt(std::max(dot(v, v), 0.0f));
    return n > 1.0e-20f ? v3(v.x * (1.0f / n), v.y * (1.0f / n), v.z * (1.0f / n)) : fallback;
}
float tree_average_abs(float x, float y, float z) { return (std::abs(x) + std::abs(y) + std::abs(z)) / 3.0f; }

TreeEmitterIn triangle_tree_emitter(uint32_t id, const Vec3 q[3], const float *o2w, const float estimate[3],
                                    uint32_t emission_sampling) {
    const Vec3 c = cross(sub(q[1], q[0]), sub(q[2], q[0]));
    const float area = 0.5f * std::sqrt(std::max(dot(c, c), 0.0f));
    const bool front = emission_sampling == PSX_EMISSION_SAMPLING_FRONT;
    const bool back = emission_sampling == PSX_EMISSION_SAMPLING_BACK;
    Vec3 axis = (front || back) ? tree_unit_or(c, v3(0.0f, 0.0f, 1.0f)) : tree_unit_or(sub(q[0], q[1]), v3(1.0f, 0.0f, 0.0f));
    if (back) axis = v3(axis.x * -1.0f, axis.y * -1.0f, axis.z * -1.0f);
    if (front || back) {
        const Vec3 a0 = v3(o2w[0], o2w[1], o2w[2]), a1 = v3(o2w[4], o2w[5], o2w[6]), a2 = v3(o2w[8], o2w[9], o2w[10]);
        if (dot(a0, cross(a1, a2)) < 0.0f) axis = v3(axis.x * -1.0f, axis.y * -1.0f, axis.z * -1.0f);
    }
    TreeEmitterIn e;
    e.measure.bounds_min[0] = std::min(std::min(q[0].x, q[1].x), q[2].x);
    e.measure.bounds_min[1] = std::min(std::min(q[0].y, q[1].y), q[2].y);
    e.measure.bounds_min[2] = std::min(std::min(q[0].z, q[1].z), q[2].z);
    e.measure.bounds_max[0] = std::max(std::max(q[0].x, q[1].x), q[2].x);
    e.measure.bounds_max[1] = std::max(std::max(q[0].y, q[1].y), q[2].y);
    e.measure.bounds_max[2] = std::max(std::max(q[0].z, q[1].z), q[2].z);
    e.measure.bounds_empty = false;
    e.measure.axis[0] = axis.x;
    e.measure.axis[1] = axis.y;
    e.measure.axis[2] = axis.z;
    e.measure.theta_o = (front || back) ? 0.0f : kTreeHalfPi;
    e.measure.theta_e = kTreeHalfPi;
    e.measure.orientation_empty = false;
    e.measure.energy = area * tree_average_abs(estimate[0], estimate[1], estimate[2]);
    const float third = 1.0f / 3.0f;
    e.centroid[0] = ((q[0].x + q[1].x) + q[2].x) * third;
    e.centroid[1] = ((q[0].y + q[1].y) + q[2].y) * third;
    e.centroid[2] = ((q[0].z + q[1].z) + q[2].z) * third;
    e.emitter_id = id;
    e.distant = false;
    return e;
}

TreeEmitterIn lamp_tree_emitter(uint32_t id, const Lamp &l) {
    float strength[3] = {l.color[0] * l.power, l.color[1] * l.power, l.color[2] * l.power};
    if ((l.flags & kLampNormalize) == 0u) {
        float area = 1.0f;
        if (l.type == kLampPoint || l.type == kLampSpot) {
            area = l.radius > 0.0f ? 4.0f * kTreePi * l.radius * l.radius : 4.0f;
        } else if (l.type == kLampArea) {
            area = std::abs(l.size_u * l.size_v);
            if ((l.flags & kLampEllipse) != 0u) area *= 0.25f * kTreePi;
        } else if (l.type == kLampDistant && l.angle > 0.0f) {
            const float s = std::sin(0.5f * l.angle);
            area = kTreePi * s * s;
        }
        for (float &v : strength) v = v * area;
    }
    TreeEmitterIn e;
    e.measure.axis[0] = l.axis_z[0] * -1.0f;
    e.measure.axis[1] = l.axis_z[1] * -1.0f;
    e.measure.axis[2] = l.axis_z[2] * -1.0f;
    e.measure.orientation_empty = false;
    e.measure.bounds_empty = true;
    for (int i = 0; i < 3; ++i) e.measure.bounds_min[i] = e.measure.bounds_max[i] = 0.0f;
    if (l.type == kLampArea) {
        e.measure.theta_e = 0.5f * l.spread;
        float hu[3], hv[3], c[4][3];
        for (int i = 0; i < 3; ++i) {
            hu[i] = l.axis_x[i] * (0.5f * l.size_u);
            hv[i] = l.axis_y[i] * (0.5f * l.size_v);
        }
        for (int i = 0; i < 3; ++i) {
            c[0][i] = l.position[i] + (hu[i] + hv[i]);
            c[1][i] = l.position[i] + (hu[i] - hv[i]);
            c[2][i] = l.position[i] + (hv[i] - hu[i]);
            c[3][i] = l.position[i] - (hu[i] + hv[i]);
        }
        for (int i = 0; i < 3; ++i) {
            e.measure.bounds_min[i] = e.measure.bounds_max[i] = c[0][i];
            for (int k = 1; k < 4; ++k) {
                e.measure.bounds_min[i] = std::min(e.measure.bounds_min[i], c[k][i]);
                e.measure.bounds_max[i] = std::max(e.measure.bounds_max[i], c[k][i]);
            }
        }
        e.measure.bounds_empty = false;
        for (float &v : strength) v = v * (1.0f / kTreePi);
    } else if (l.type == kLampPoint) {
        e.measure.theta_o = kTreePi;
        e.measure.theta_e = kTreeHalfPi;
    } else if (l.type == kLampSpot) {
        float theta_e = std::min(0.5f * l.spot_angle, kTreeHalfPi);
        if (std::abs(kTreeHalfPi - theta_e) >= 1.0e-6f)
            theta_e = std::atan(std::tan(theta_e) * std::max(l.axis_scale[0], l.axis_scale[1]) /
                                std::max(l.axis_scale[2], 1.0e-20f));
        e.measure.theta_e = theta_e;
    } else {
        e.measure.theta_e = 0.5f * l.angle;
        e.distant = true;
    }
    if (l.type == kLampPoint || l.type == kLampSpot) {
        for (int i = 0; i < 3; ++i) {
            const float a = l.position[i] - l.radius, b = l.position[i] + l.radius;
            e.measure.bounds_min[i] = std::min(a, b);
            e.measure.bounds_max[i] = std::max(a, b);
        }
        e.measure.bounds_empty = false;
        const float k = 0.25f / kTreePi;
        for (float &v : strength) v = v * k;
    }
    e.measure.energy = tree_average_abs(strength[0], strength[1], strength[2]);
    for (int i = 0; i < 3; ++i) e.centroid[i] = l.position[i];
    e.emitter_id = id;
    return e;
}

// emissive triangles (src/luisa/path_tracer_scene.cpp:1548-1640)
void Scene::index_emissive_candidates() {
    std::vector<bool> can_emit(materials.size(), false);
    for (size_t m = 0; m < materials.size(); ++m) {
        bool emission = false;
        for (const psx_closure_instr &ins : materials[m].instr) emission |= ins.op == PSX_CLOSURE_EMISSION;
        can_emit[m] = emission && materials[m].emission_sampling != PSX_EMISSION_SAMPLING_NONE;
    }
    emissive_candidates.clear();
    for (uint32_t i = 0; i < instances.size(); ++i) {
        const psx_instance &inst = instances[i];
        const bool light_visible =
            inst.visibility_mask == 0xffffffffu ||
            (inst.visibility_mask & (PSX_VISIBILITY_DIFFUSE | PSX_VISIBILITY_GLOSSY | PSX_VISIBILITY_TRANSMISSION |
                                     PSX_VISIBILITY_VOLUME_SCATTER)) != 0u;
        if (!light_visible) continue;
        const uint32_t mi = inst.mesh;
        const Mesh &mesh = meshes[mi];
        for (uint32_t p = 0; p < mesh.triangle_count; ++p) {
            const uint32_t slot = mesh_tri_slots[mi].empty() ? 0u : mesh_tri_slots[mi][p];
            uint32_t mat = kNone;
            if (slot < overrides[i].size())
                mat = overrides[i][slot];
            else if (!mesh_slots[mi].empty())
                mat = mesh_slots[mi][std::min<size_t>(slot, mesh_slots[mi].size() - 1u)];
            if (mat >= materials.size() || !can_emit[mat]) continue;
            emissive_candidates.push_back(EmissiveCandidate{i, p, mat});
        }
    }
    candidates_indexed = true;
}

void Scene::rebuild_emitters(ParamTables &t) {
    t.light = std::make_shared<EmitterTables>(table_stream);
    EmitterTables &lt = *t.light;
    std::vector<Emitter> em;
    std::vector<float> areas;
    std::vector<TreeEmitterIn> tree_in;
    std::vector<uint32_t> cprims;
    if (!candidates_indexed) index_emissive_candidates();
    {
        for (const EmissiveCandidate &cand : emissive_candidates) {
            const uint32_t i = cand.instance, p = cand.prim, mat = cand.material;
            if (!materials[mat].may_emit) continue;
            const psx_instance &inst = instances[i];
            const uint32_t mi = inst.mesh;
            const auto &pos = mesh_positions[mi];
            const auto &tri = mesh_tris[mi];
            Vec3 q[3];
            for (int c = 0; c < 3; ++c) {
                const uint32_t vi = tri[3u * p + c];
                q[c] = transform_point(inst.transform, v3(pos[3u * vi], pos[3u * vi + 1], pos[3u * vi + 2]));
            }
            const Vec3 n = cross(sub(q[1], q[0]), sub(q[2], q[0]));
            areas.push_back(0.5f * std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z));
            float estimate[3] = {0.0f, 0.0f, 0.0f};
            if (!materials[mat].instr.empty()) emission_estimate(materials[mat], materials[mat].root, estimate, 0);
            tree_in.push_back(triangle_tree_emitter(static_cast<uint32_t>(em.size()), q, inst.transform, estimate,
                                                    materials[mat].emission_sampling));
            cprims.push_back(mesh_prim_offsets[mi] + p);
            em.push_back(Emitter{i, p, mat, materials[mat].emission_sampling});
        }
    }
    for (size_t j = 0; j < lamp_list.size(); ++j)
        tree_in.push_back(lamp_tree_emitter(static_cast<uint32_t>(em.size() + j), lamp_list[j]));
    // light tree over triangles, then lamps
    // (src/luisa/path_tracer_light_sampling_scene.cpp:150-245)
    LightTreeTables tree;
    if (!build_light_tree(tree_in, 8u, tree) || !tree.usable) tree = LightTreeTables{};
    lt.tree_nodes.upload(tree.nodes);
    lt.tree_emitters.upload(tree.emitters);
    lt.tree_mapping.upload(tree.mapping);
    lt.emitter_cprim.upload(cprims);
    view.tree_nodes = lt.tree_nodes.as<TreeNode>();
    view.tree_emitters = lt.tree_emitters.as<TreeEmitter>();
    view.tree_mapping = lt.tree_mapping.as<uint2>();
    view.emitter_cprim = lt.emitter_cprim.as<uint32_t>();
    view.tree_node_count = static_cast<uint32_t>(tree.nodes.size());
    view.tree_emitter_count = static_cast<uint32_t>(tree.emitters.size());
    view.tree_root = tree.root;
    std::vector<DistEntry> dist;
    float tri_pdf = 0.0f;
    float lamp_pdf = 0.0f;
    const uint32_t count = build_distribution(areas, static_cast<uint32_t>(lamp_list.size()), dist, tri_pdf, lamp_pdf);
    lt.emitters.upload(em);
    lt.distribution.upload(dist);
    view.light_selection_pdf = lamp_pdf;
    view.emitters = lt.emitters.as<Emitter>();
    view.distribution = lt.distribution.as<DistEntry>();
    view.emitter_count = static_cast<uint32_t>(em.size());
    view.distribution_count = count;
    view.triangle_area_pdf = tri_pdf;
}

// Wavefront path state of one session (psx_device.h, WavefrontView), sized
// for max_samples_per_dispatch samples of every window pixel.
struct WavefrontState {
    DeviceBuffer ray_o, ray_d, aux, throughput, mis_normal, radiance, normal, albedo, state0, state1, ident, hit, hit_uv,
        shadow_o, shadow_d, shadow_q, shadow_l, shadow_t, queue0, queue1, list0, list1, counters, lp, shadow_split;
    WavefrontView view{};

    bool light = false;  // light-pass arrays allocated

    void allocate(uint32_t capacity, uint32_t padded_pixels, uint32_t iterations, bool light_passes) {
        const size_t n = capacity;
        light = light_passes;
        if (light_passes) {
            lp.allocate(n * 10u * 16u);
            shadow_split.allocate(n * 2u * 16u);
        } else {
            lp.reset();
            shadow_split.reset();
        }
        ray_o.allocate(n * 16u);
        ray_d.allocate(n * 16u);
        aux.allocate(n * 16u);
        throughput.allocate(n * 16u);
        mis_normal.allocate(n * 16u);
        radiance.allocate(n * 16u);
        normal.allocate(n * 16u);
        albedo.allocate(n * 16u);
        state0.allocate(n * 16u);
        state1.allocate(n * 16u);
        ident.allocate(n * 8u);
        hit.allocate(n * 16u);
        hit_uv.allocate(n * 8u);
        shadow_o.allocate(n * 16u);
        shadow_d.allocate(n * 16u);
        shadow_q.allocate(n * 16u);
        shadow_l.allocate(n * 16u);
        shadow_t.allocate(n * 16u);
        // a live path queues at most a continuation ray and a shadow ray;
        // shade waves leave partly used reservations as skip entries
        const size_t qh = wavefront_queue_headroom(), lh = wavefront_list_headroom();
        queue0.allocate((2u * n + qh) * 4u);
        queue1.allocate((2u * n + qh) * 4u);
        list0.allocate((n + lh) * 4u);
        list1.allocate((n + lh) * 4u);
        counters.allocate((iterations + 1u) * sizeof(WavefrontCounters));
        WavefrontView &v = view;
        v.ray_o = ray_o.as<float4>();
        v.ray_d = ray_d.as<float4>();
        v.aux = aux.as<uint4>();
        v.throughput = throughput.as<float4>();
        v.mis_normal = mis_normal.as<float4>();
        v.radiance = radiance.as<float4>();
        v.normal = normal.as<float4>();
        v.albedo = albedo.as<float4>();
        v.state0 = state0.as<uint4>();
        v.state1 = state1.as<uint4>();
        v.ident = ident.as<uint2>();
        v.hit = hit.as<uint4>();
        v.hit_uv = hit_uv.as<float2>();
        v.shadow_o = shadow_o.as<float4>();
        v.shadow_d = shadow_d.as<float4>();
        v.shadow_q = shadow_q.as<uint4>();
        v.shadow_l = shadow_l.as<float4>();
        v.shadow_t = shadow_t.as<float4>();
        v.queue[0] = queue0.as<uint32_t>();
        v.queue[1] = queue1.as<uint32_t>();
        v.list[0] = list0.as<uint32_t>();
        v.list[1] = list1.as<uint32_t>();
        v.counters = counters.as<WavefrontCounters>();
        v.lp = light_passes ? lp.as<float4>() : nullptr;
        v.shadow_split = light_passes ? shadow_split.as<float4>() : nullptr;
        v.capacity = capacity;
        v.padded_pixels = padded_pixels;
        psx_wavefront_tuning t;
        psx_wavefront_tuning_default(&t);
        set_tuning(t);
    }

    // launch shape (a valid psx_wavefront_tuning) into the view
    void set_tuning(const psx_wavefront_tuning &t) {
        WavefrontView &v = view;
        v.tile_shift = static_cast<uint32_t>(__builtin_ctz(t.tile_width));
        v.refill = t.refill_lanes;
        v.trace_chunk = t.trace_chunk;
        v.queue_chunk = t.queue_chunk;
        v.list_chunk = t.list_chunk;
        v.shade_sort = t.sort_shading;
        v.segments = t.xcd_segments;
    }
};

bool valid_tuning(const psx_wavefront_tuning &t) {
    const bool tile = t.tile_width == 8u || t.tile_width == 16u || t.tile_width == 32u || t.tile_width == 64u;
    const bool list = t.list_chunk == 64u || t.list_chunk == 128u || t.list_chunk == 256u;
    return tile && t.refill_lanes >= 1u && t.refill_lanes <= 64u && t.trace_chunk >= 16u && t.trace_chunk <= 4096u &&
           (t.queue_chunk == 128u || t.queue_chunk == 256u) && list && t.sort_shading <= 1u &&
           (t.xcd_segments == 1u || t.xcd_segments == 8u);
}

// A wavefront dispatch captured as a graph, replayed while every kernel
// argument it was captured with (`key`) is unchanged.
struct DispatchGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;

    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    ~DispatchGraph() { reset(); }
};

struct Session {
    Scene *scene = nullptr;
    psx_render_settings settings{};
    std::vector<uint32_t> passes, pass_channels;
    KernelParams params{};
    uint32_t window_w = 0, window_h = 0;
    DeviceBuffer combined, normal, albedo, sample_count, sobol, filter, counters, spill, tile_cursor;
    DeviceBuffer light;        // light-pass sums (kLightPasses float4 per pixel) when requested
    bool light_passes = false;
    WavefrontState wavefront;
    uint32_t wavefront_iterations = 0;
    uint32_t sobol_size = 0;
    uint32_t total_aa = 0;
    uint32_t rendered = 0;
    std::atomic<bool> cancelled{false};
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    std::vector<hipEvent_t> trace_events;  // pairs around wavefront trace launches
    psx_render_stats stats{};
    bool count_nodes = false;
    psx_wavefront_tuning tuning{8u, 16u, 128u, 256u, 128u, 0u, 1u};
    DispatchGraph graph[2];       // full dispatches of max_samples_per_dispatch, and the remainder
    bool graph_failed = false;    // capture failed once: launch from the host
    DeviceBuffer dispatch_first;  // first sample of the dispatch a replayed graph renders

    ~Session() {
        graph[0].reset();
        graph[1].reset();
        for (hipEvent_t e : trace_events) (void)hipEventDestroy(e);
        if (ev_begin) (void)hipEventDestroy(ev_begin);
        if (ev_end) (void)hipEventDestroy(ev_end);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// LightPassBuffer slot of a pass kind, or -1
// (src/luisa/path_tracer_internal.h:410-424)
int light_pass_slot(uint32_t kind) {
    switch (kind) {
    case PSX_PASS_DIFFUSE_DIRECT: return kLpDiffuseDirect;
    case PSX_PASS_DIFFUSE_INDIRECT: return kLpDiffuseIndirect;
    case PSX_PASS_GLOSSY_DIRECT: return kLpGlossyDirect;
    case PSX_PASS_GLOSSY_INDIRECT: return kLpGlossyIndirect;
    case PSX_PASS_TRANSMISSION_DIRECT: return kLpTransmissionDirect;
    case PSX_PASS_TRANSMISSION_INDIRECT: return kLpTransmissionIndirect;
    case PSX_PASS_VOLUME_DIRECT: return kLpVolumeDirect;
    case PSX_PASS_VOLUME_INDIRECT: return kLpVolumeIndirect;
    case PSX_PASS_EMISSION: return kLpEmission;
    case PSX_PASS_ENVIRONMENT: return kLpEnvironment;
    case PSX_PASS_GLOSSY_COLOR: return kLpGlossyColor;
    case PSX_PASS_TRANSMISSION_COLOR: return kLpTransmissionColor;
    default: return -1;
    }
}

// Cycles' safe_divide_even_color (src/luisa/path_tracer_common.cpp:516-548):
// channels with a zero denominator take the other channels' quotients.
void divide_even_color(const float *n, const float *d, float *out) {
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

// Kernel parameters of one render request: the host part of
// LuisaRenderSession::initialize (src/luisa/path_tracer_kernel.cpp:44-439).
// `view` is the scene version the request renders (NEE needs a non-empty
// light distribution).
KernelParams kernel_params(const Scene &scene, const SceneView &view, const psx_render_settings &rs) {
    KernelParams k{};
    k.window_x = rs.window_x;
    k.window_y = rs.window_y;
    k.window_width = rs.window_width ? rs.window_width : rs.width;
    k.window_height = rs.window_height ? rs.window_height : rs.height;
    k.full_width = std::max(rs.width, 1u);
    k.full_height = std::max(rs.height, 1u);
    k.seed = rs.seed;
    const psx_integrator_settings &in = rs.integrator;
    auto synced = [](uint32_t v) { return v == 0xffffffffu ? v : v + 1u; };
    k.max_bounces = synced(in.max_bounces);
    k.min_bounces = synced(in.min_bounces);
    k.max_diffuse = synced(in.diffuse_bounces);
    k.max_glossy = synced(in.glossy_bounces);
    k.max_transmission = synced(in.transmission_bounces);
    k.transparent_min = synced(in.transparent_min_bounces);
    k.transparent_max = in.transparent_max_bounces;
    const uint64_t req = 2ull * k.max_bounces + std::max(in.transparent_max_bounces, 1u) + 1ull;
    k.max_path_steps = static_cast<uint32_t>(std::min<uint64_t>(req, 1024ull));
    k.transparent_background = rs.transparent_background;
    k.clamp_direct = in.sample_clamp_direct > 0.0f ? in.sample_clamp_direct * 3.0f : 0.0f;
    k.clamp_indirect = in.sample_clamp_indirect > 0.0f ? in.sample_clamp_indirect * 3.0f : 0.0f;
    k.light_inv_rr_threshold =
        (!in.use_light_tree && in.light_sampling_threshold > 0.0f) ? in.film_exposure / in.light_sampling_threshold
                                                                   : 0.0f;
    const bool available = view.distribution_count > 0u;
    const bool nee = scene.backend->options.next_event_estimation != 0u &&
                     in.direct_light_sampling != PSX_DIRECT_LIGHT_FORWARD && available;
    k.nee = nee ? 1u : 0u;
    k.use_tree = in.use_light_tree ? 1u : 0u;
    k.mis_mode = nee ? in.direct_light_sampling : PSX_DIRECT_LIGHT_FORWARD;
    const psx_camera &cam = scene.camera;
    const float aspect = static_cast<float>(k.full_width) / static_cast<float>(k.full_height);
    const bool hfit = cam.sensor_fit == 0u || (cam.sensor_fit == 2u && aspect >= 1.0f);
    k.tan_v = hfit ? std::tan(cam.horizontal_field_of_view * 0.5f) / aspect : std::tan(cam.field_of_view * 0.5f);
    k.tan_h = k.tan_v * aspect;
    // orthographic_viewplane_span (include/psycles/luisa/camera_sampling.h:50-56)
    k.projection = cam.projection;
    k.ortho_span = hfit ? cam.orthographic_scale / aspect : cam.orthographic_scale;
    k.shift_x = cam.lens_shift_x * (hfit ? 1.0f : 1.0f / aspect);
    k.shift_y = cam.lens_shift_y * (hfit ? aspect : 1.0f);
    k.near_clip = cam.near_clip;
    k.far_clip = cam.far_clip;
    k.pass_alpha_threshold = std::min(std::max(rs.pass_alpha_threshold, 0.0f), 1.0f);
    std::memcpy(k.camera, cam.transform, sizeof(k.camera));
    return k;
}

// Copies the settings (with their pass arrays) into the session.
void keep_settings(Session &s, const psx_render_settings &rs) {
    s.settings = rs;
    s.passes.assign(rs.passes, rs.passes + rs.pass_count);
    if (rs.pass_channels)
        s.pass_channels.assign(rs.pass_channels, rs.pass_channels + rs.pass_count);
    else
        s.pass_channels.assign(rs.pass_count, 4u);
    s.settings.passes = s.passes.data();
    s.settings.pass_channels = s.pass_channels.data();
}

// LuisaRenderSession::initialize (src/luisa/path_tracer_kernel.cpp:44-439)
void initialize_session(Session &s, const psx_render_settings &rs) {
    keep_settings(s, rs);
    s.params = kernel_params(*s.scene, s.scene->view, rs);
    const KernelParams &k = s.params;
    s.window_w = k.window_width;
    s.window_h = k.window_height;
    const size_t pixels = static_cast<size_t>(s.window_w) * s.window_h;
    if (s.sample_count.bytes() != std::max<size_t>(pixels * 4u, 16u)) {
        s.combined.allocate(pixels * 16u);
        s.normal.allocate(pixels * 16u);
        s.albedo.allocate(pixels * 16u);
        s.sample_count.allocate(pixels * 4u);
    }
    hip_check(hipMemset(s.combined.as<void>(), 0, pixels * 16u));
    hip_check(hipMemset(s.normal.as<void>(), 0, pixels * 16u));
    hip_check(hipMemset(s.albedo.as<void>(), 0, pixels * 16u));
    hip_check(hipMemset(s.sample_count.as<void>(), 0, pixels * 4u));
    s.light_passes = false;
    for (uint32_t kind : s.passes) s.light_passes |= light_pass_slot(kind) >= 0;
    if (s.light_passes) {
        if (s.light.bytes() != pixels * kLightPasses * 16u) s.light.allocate(pixels * kLightPasses * 16u);
        hip_check(hipMemset(s.light.as<void>(), 0, pixels * kLightPasses * 16u));
    } else {
        s.light.reset();
    }
    s.filter.upload(pixel_filter_table(rs.pixel_filter, rs.filter_width));
    if (!s.counters.bytes()) s.counters.allocate(sizeof(Counters));
    s.total_aa = 0;
    s.rendered = 0;
    s.sobol_size = 0;
}

// prepare_sobol_table (src/luisa/path_tracer_session.cpp:41-69)
void prepare_sobol(Session &s, uint32_t total) {
    const uint32_t size = sequence_size_for(total);
    if (size != s.sobol_size) {
        s.sobol.upload(sobol_table(size));
        s.sobol_size = size;
    }
    s.params.sequence_size = size;
}

bool valid_settings(const psx_render_settings *rs) {
    if (!rs || rs->width == 0u || rs->height == 0u) return false;
    const uint32_t ww = rs->window_width ? rs->window_width : rs->width;
    const uint32_t wh = rs->window_height ? rs->window_height : rs->height;
    if (static_cast<uint64_t>(rs->window_x) + ww > rs->width) return false;
    if (static_cast<uint64_t>(rs->window_y) + wh > rs->height) return false;
    if (rs->pass_count && !rs->passes) return false;
    for (uint32_t i = 0; i < rs->pass_count; ++i) {
        const uint32_t p = rs->passes[i];
        if (p != PSX_PASS_COMBINED && p != PSX_PASS_NORMAL && p != PSX_PASS_ALBEDO && p != PSX_PASS_SAMPLE_COUNT &&
            p != PSX_PASS_DENOISING_NORMAL && p != PSX_PASS_DENOISING_ALBEDO && light_pass_slot(p) < 0)
            return false;
    }
    return true;
}

// Runs samples [first, first + count) as ordered batches of at most
// max_samples_per_dispatch (SampleDispatchPartition::make,
// src/luisa/sample_dispatch_partition.h:134-235).
// Wavefront buffers for `per` samples per dispatch; kept across dispatches
// and sessions resets while the window and limits fit.
void prepare_wavefront(Session &s, uint32_t per) {
    const uint32_t shift = static_cast<uint32_t>(__builtin_ctz(s.tuning.tile_width));
    const uint32_t padded = wavefront_padded_pixels(s.params, shift);
    const uint64_t want = static_cast<uint64_t>(padded) * per;
    if (want > 0x7fffffffull) throw DeviceError{PSX_ERROR_INVALID_ARGUMENT};
    const uint32_t iterations = s.params.max_path_steps + 1u;
    const WavefrontView &v = s.wavefront.view;
    if (v.capacity < want || s.wavefront_iterations < iterations || s.wavefront.light != s.light_passes) {
        s.wavefront.allocate(static_cast<uint32_t>(want), padded, iterations, s.light_passes);
        s.wavefront_iterations = iterations;
    }
    s.wavefront.view.padded_pixels = padded;
    s.wavefront.set_tuning(s.tuning);
    if (!s.dispatch_first.bytes()) s.dispatch_first.allocate(sizeof(uint32_t));
    const size_t events = 2u * static_cast<size_t>(iterations);
    while (s.trace_events.size() < events) {
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e));
        s.trace_events.push_back(e);
    }
}

// The scene's generated shade kernel for `variant`, compiled on first use.
GeneratedShade generated_program(Scene &scene, uint32_t variant) {
    try {
        return scene.backend->programs.get(scene.programs, variant, scene.backend->device);
    } catch (const std::runtime_error &e) {
        g_last_error = e.what();
        throw DeviceError{PSX_ERROR_DEVICE};
    }
}

// appends the bytes of `v` to a graph key
template <class T>
void key_bytes(std::vector<unsigned char> &key, const T &v) {
    const auto *b = reinterpret_cast<const unsigned char *>(&v);
    key.insert(key.end(), b, b + sizeof(T));
}

// Captures one wavefront dispatch of `count` samples (every kernel launch of
// launch_wavefront) into a graph slot.  The dispatch
// reads its first sample from `dispatch_first`, so one graph serves every
// range of the same size.  Returns false, with no graph, when the capture
// or instantiation fails; the session then launches from the host.
bool capture_dispatch(Session &s, DispatchGraph &slot, const FilmView &film, uint32_t count,
                      const GeneratedShade *generated) {
    slot.reset();
    hipGraph_t g = nullptr;
    std::lock_guard<std::mutex> gate(s.scene->backend->capture_gate);
    if (hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
    const hipError_t launched =
        launch_wavefront(s.scene->view, s.params, s.sobol.as<float4>(), s.filter.as<float>(), film, s.wavefront.view,
                         0u, count, s.dispatch_first.as<uint32_t>(), s.counters.as<Counters>(), s.count_nodes,
                         s.spill.as<uint32_t>(), nullptr, generated, s.stream);
    const hipError_t ended = hipStreamEndCapture(s.stream, &g);
    hipGraphExec_t exec = nullptr;
    const bool ok = launched == hipSuccess && ended == hipSuccess && g &&
                    hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (!ok) return false;
    slot.exec = exec;
    return true;
}

int dispatch_samples(Session &s, uint32_t first, uint32_t count) {
    const psx_backend_options &opt = s.scene->backend->options;
    const uint32_t per = std::max(opt.max_samples_per_dispatch, 1u);
    // light passes are accumulated by the wavefront kernels (both integrators
    // give bit-identical films, so the choice is invisible in the result)
    const bool wavefront = opt.integrator == PSX_INTEGRATOR_WAVEFRONT || s.light_passes;
    if (wavefront) prepare_wavefront(s, std::min(per, count));
    hip_check(hipMemsetAsync(s.counters.as<void>(), 0, sizeof(Counters), s.stream));
    FilmView film{s.combined.as<float>(), s.normal.as<float>(), s.albedo.as<float>(), s.sample_count.as<uint32_t>(),
                  s.light_passes ? s.light.as<float>() : nullptr};
    hip_check(hipEventRecord(s.ev_begin, s.stream));
    uint64_t dispatches = 0, graph_launches = 0, graph_builds = 0;
    // events time every trace launch; they cannot be timed inside a graph
    // (hipEventElapsedTime rejects events recorded by graph nodes), so timed
    // dispatches launch from the host
    const bool time_trace = (opt.flags & PSX_BACKEND_TIME_TRACE) != 0u;
    const bool graphs = (opt.flags & PSX_BACKEND_NO_GRAPHS) == 0u && !time_trace && !s.graph_failed;
    GeneratedShade generated_shade{};
    const GeneratedShade *generated = nullptr;
    if (wavefront && (opt.flags & PSX_BACKEND_GENERATED_PROGRAMS) != 0u) {
        generated_shade = generated_program(*s.scene, shade_variant(s.scene->view, s.light_passes));
        generated = &generated_shade;
    }
    double trace_ms = 0.0;
    uint64_t trace_launches = 0;
    for (uint32_t done = 0; done < count;) {
        if (s.cancelled.load()) {
            hip_check(hipStreamSynchronize(s.stream));
            return PSX_ERROR_CANCELLED;
        }
        const uint32_t n = std::min(per, count - done);
        if (wavefront) {
            bool replayed = false;
            if (graphs) {
                std::vector<unsigned char> key;
                key_bytes(key, s.scene->view);
                key_bytes(key, s.params);
                key_bytes(key, film);
                key_bytes(key, s.wavefront.view);
                const void *ptrs[4] = {s.sobol.as<void>(), s.filter.as<void>(), s.counters.as<void>(),
                                       s.spill.as<void>()};
                key_bytes(key, ptrs);
                key_bytes(key, n);
                key_bytes(key, s.count_nodes);
                // members one by one: the struct's tail padding is not initialised
                key_bytes(key, generated_shade.fn);
                key_bytes(key, generated_shade.blocks);
                DispatchGraph &slot = s.graph[n == per ? 0 : 1];
                if (!slot.exec || key != slot.key) {
                    if (capture_dispatch(s, slot, film, n, generated)) {
                        slot.key = std::move(key);
                        ++graph_builds;
                    } else {
                        s.graph_failed = true;
                    }
                }
                if (slot.exec) {
                    hip_check(launch_store_word(s.dispatch_first.as<uint32_t>(), first + done, s.stream));
                    hip_check(hipGraphLaunch(slot.exec, s.stream));
                    ++graph_launches;
                    replayed = true;
                }
            }
            if (!replayed)
                hip_check(launch_wavefront(s.scene->view, s.params, s.sobol.as<float4>(), s.filter.as<float>(), film,
                                           s.wavefront.view, first + done, n, nullptr, s.counters.as<Counters>(),
                                           s.count_nodes, s.spill.as<uint32_t>(),
                                           time_trace ? s.trace_events.data() : nullptr, generated, s.stream));
            if (time_trace) {
                // the event pool is reused by the next dispatch: read it now
                hip_check(hipStreamSynchronize(s.stream));
                for (uint32_t it = 0; it < s.params.max_path_steps + 1u; ++it) {
                    float ms = 0.0f;
                    hip_check(hipEventElapsedTime(&ms, s.trace_events[2u * it], s.trace_events[2u * it + 1u]));
                    trace_ms += ms;
                    ++trace_launches;
                }
            }
        } else {
            hip_check(launch_render(s.scene->view, s.params, s.sobol.as<float4>(), s.filter.as<float>(), film,
                                    first + done, n, s.counters.as<Counters>(), s.count_nodes, s.spill.as<uint32_t>(),
                                    s.tile_cursor.as<uint32_t>(), s.stream));
        }
        ++dispatches;
        done += n;
    }
    hip_check(hipEventRecord(s.ev_end, s.stream));
    hip_check(hipStreamSynchronize(s.stream));
    Counters c{};
    hip_check(hipMemcpy(&c, s.counters.as<void>(), sizeof(c), hipMemcpyDeviceToHost));
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, s.ev_begin, s.ev_end));
    s.stats.rays = c.rays;
    s.stats.node_visits = c.nodes;
    s.stats.triangle_tests = c.tris;
    s.stats.samples = static_cast<uint64_t>(count) * s.window_w * s.window_h;
    s.stats.dispatches = dispatches;
    s.stats.kernel_ms = ms;
    s.stats.trace_ms = trace_ms;
    s.stats.trace_launches = trace_launches;
    s.stats.graph_launches = graph_launches;
    s.stats.graph_builds = graph_builds;
    s.rendered += count;
    return PSX_OK;
}

// write_passes (src/luisa/path_tracer_session.cpp:71-254)
int write_passes(Session &s, const psx_sink *sink) {
    const size_t n = static_cast<size_t>(s.window_w) * s.window_h;
    std::vector<float> combined(n * 4u), normal(n * 4u), albedo(n * 4u);
    std::vector<uint32_t> counts(n);
    hip_check(hipMemcpy(combined.data(), s.combined.as<void>(), n * 16u, hipMemcpyDeviceToHost));
    hip_check(hipMemcpy(normal.data(), s.normal.as<void>(), n * 16u, hipMemcpyDeviceToHost));
    hip_check(hipMemcpy(albedo.data(), s.albedo.as<void>(), n * 16u, hipMemcpyDeviceToHost));
    hip_check(hipMemcpy(counts.data(), s.sample_count.as<void>(), n * 4u, hipMemcpyDeviceToHost));
    std::vector<float> light;
    if (s.light_passes) {
        light.resize(n * kLightPasses * 4u);
        hip_check(hipMemcpy(light.data(), s.light.as<void>(), n * kLightPasses * 16u, hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < n; ++i)
        if (counts[i] != s.rendered) return PSX_ERROR_INCOMPLETE;
    if (sink && sink->begin) sink->begin(sink->user, &s.settings);
    const float exposure = s.settings.integrator.film_exposure;
    std::vector<float> pixels;
    for (uint32_t p = 0; p < s.passes.size(); ++p) {
        const uint32_t kind = s.passes[p];
        const uint32_t channels = std::max(1u, std::min(s.pass_channels[p], 4u));
        pixels.assign(n * channels, 0.0f);
        for (size_t i = 0; i < n; ++i) {
            const float den = static_cast<float>(std::max(counts[i], 1u));
            float v[4] = {0, 0, 0, 0};
            switch (kind) {
            case PSX_PASS_COMBINED:
                for (int c = 0; c < 3; ++c) v[c] = combined[4 * i + c] * (exposure / den);
                v[3] = std::min(std::max(1.0f - combined[4 * i + 3] / den, 0.0f), 1.0f);
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
                v[0] = v[1] = v[2] = v[3] = static_cast<float>(counts[i]);
                break;
            default: {
                // light passes (src/luisa/path_tracer_session.cpp:150-230)
                const int slot = light_pass_slot(kind);
                if (slot < 0 || light.empty()) break;
                const float *lp = light.data() + i * kLightPasses * 4u;
                const float *sum = lp + 4 * slot;
                const float *color = nullptr;
                if (slot == kLpDiffuseDirect || slot == kLpDiffuseIndirect) color = albedo.data() + 4 * i;
                if (slot == kLpGlossyDirect || slot == kLpGlossyIndirect) color = lp + 4 * kLpGlossyColor;
                if (slot == kLpTransmissionDirect || slot == kLpTransmissionIndirect)
                    color = lp + 4 * kLpTransmissionColor;
                if (color) {
                    divide_even_color(sum, color, v);
                    for (int c = 0; c < 3; ++c) v[c] *= exposure;
                    v[3] = 1.0f;
                } else if (slot == kLpGlossyColor || slot == kLpTransmissionColor) {
                    for (int c = 0; c < 4; ++c) v[c] = sum[c] / den;
                } else {
                    for (int c = 0; c < 4; ++c) v[c] = sum[c] * (exposure / den);
                }
                break;
            }
            }
            for (uint32_t c = 0; c < channels; ++c) pixels[i * channels + c] = v[std::min(c, 3u)];
        }
        if (sink && sink->write)
            sink->write(sink->user, kind, channels, s.settings.window_x, s.settings.window_y, s.window_w, s.window_h,
                        s.settings.width, s.settings.height, pixels.data());
    }
    if (sink && sink->end) sink->end(sink->user, s.cancelled.load() ? 1 : 0);
    return PSX_OK;
}

// Scene compilation (LuisaPathTracerBackend::compile_scene,
// src/luisa/path_tracer_scene.cpp:50-1995, triangle subset)
// Children per BVH node the backend builds with (PSX_BACKEND_BVH_WIDTH_*).
uint32_t bvh_width(const psx_backend_options &o) {
    if (o.flags & PSX_BACKEND_BVH_WIDTH_2) return 2u;
    if (o.flags & PSX_BACKEND_BVH_WIDTH_3) return 3u;
    return 4u;
}

int compile(Backend *backend, const psx_scene_desc *d, psx_diag_fn diag, void *diag_user, Scene **out) {
    auto fail = [&](const std::string &msg) {
        if (diag) diag(diag_user, msg.c_str());
        return PSX_ERROR_INVALID_ARGUMENT;
    };
    if (d->abi_version != PSX_ABI_VERSION) return fail("scene ABI version mismatch");
    if (d->camera.projection > 1u) return fail("only perspective and orthographic cameras are supported");
    auto scene = std::make_unique<Scene>();
    scene->backend = backend;
    scene->camera = d->camera;
    // materials
    std::vector<Leaf> leaves;
    std::vector<float> params;
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const psx_material &src = d->materials[i];
        if (src.instruction_count && !src.instructions) return fail("material without instructions");
        if (src.parameter_count && !src.parameters) return fail("material without parameters");
        HostMaterial m = host_material(src);
        m.signature = structure_signature(m);
        Material g{};
        g.leaf_offset = static_cast<uint32_t>(leaves.size());
        g.param_offset = static_cast<uint32_t>(params.size());
        {
            std::vector<Leaf> ml;
            std::string error;
            if (!lower_leaves(m, ml, error)) return fail("material " + std::to_string(i) + ": " + error);
            leaves.insert(leaves.end(), ml.begin(), ml.end());
            g.leaf_count = static_cast<uint32_t>(ml.size());
            g.program = program_index(scene->programs, ml);
        }
        params.insert(params.end(), m.params.begin(), m.params.end());
        refresh_may_emit(m);
        g.emission_sampling = m.emission_sampling;
        g.may_emit = m.may_emit ? 1u : 0u;
        scene->gpu_materials.push_back(g);
        scene->materials.push_back(std::move(m));
    }
    // shade sort key per material: one per closure structure, in order of
    // first use (structures beyond the first kShadeKeys - 2 share the last key)
    std::vector<uint32_t> material_class(d->material_count);
    {
        std::vector<uint64_t> seen;
        for (uint32_t i = 0; i < d->material_count; ++i) {
            const uint64_t sig = scene->materials[i].signature;
            const auto it = std::find(seen.begin(), seen.end(), sig);
            if (it == seen.end()) seen.push_back(sig);
            const uint32_t c = static_cast<uint32_t>(it == seen.end() ? seen.size() : (it - seen.begin()) + 1);
            material_class[i] = std::min<uint32_t>(c, kShadeKeys - 1u);
        }
    }
    // meshes
    std::vector<float> positions;
    std::vector<uint32_t> indices, tri_slots, slots;
    std::vector<float> normals;
    std::vector<uint8_t> smooth;
    uint32_t prim_offset = 0;
    std::vector<uint32_t> mesh_prim_offset;
    std::vector<uint32_t> users(d->mesh_count, 0u);
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        if (d->instances[i].mesh >= d->mesh_count) return fail("instance references a missing mesh");
        users[d->instances[i].mesh]++;
    }
    for (uint32_t i = 0; i < d->mesh_count; ++i) {
        const psx_mesh &src = d->meshes[i];
        if ((src.vertex_count && !src.positions) || (src.triangle_count && !src.triangles))
            return fail("mesh " + std::to_string(i) + " without geometry arrays");
        for (uint64_t t = 0; t < 3ull * src.triangle_count; ++t)
            if (src.triangles[t] >= src.vertex_count)
                return fail("mesh " + std::to_string(i) + " has out-of-range vertex indices");
        for (uint32_t s = 0; s < src.material_slot_count; ++s)
            if (src.material_slots[s] >= d->material_count)
                return fail("mesh " + std::to_string(i) + " references a missing material");
        Mesh m{};
        m.vertex_offset = static_cast<uint32_t>(positions.size() / 3u);
        m.index_offset = static_cast<uint32_t>(indices.size() / 3u);
        m.triangle_count = src.triangle_count;
        positions.insert(positions.end(), src.positions, src.positions + 3ull * src.vertex_count);
        indices.insert(indices.end(), src.triangles, src.triangles + 3ull * src.triangle_count);
        m.normal_offset = kNone;
        if (src.normals) {
            m.normal_offset = static_cast<uint32_t>(normals.size() / 3u);
            normals.insert(normals.end(), src.normals, src.normals + 3ull * src.vertex_count);
        }
        m.smooth_offset = kNone;
        if (src.triangle_smooth) {
            m.smooth_offset = static_cast<uint32_t>(smooth.size());
            smooth.insert(smooth.end(), src.triangle_smooth, src.triangle_smooth + src.triangle_count);
        }
        m.slot_offset = static_cast<uint32_t>(slots.size());
        m.slot_count = src.material_slot_count;
        slots.insert(slots.end(), src.material_slots, src.material_slots + src.material_slot_count);
        m.tri_slot_offset = kNone;
        if (src.triangle_material_slots) {
            m.tri_slot_offset = static_cast<uint32_t>(tri_slots.size());
            tri_slots.insert(tri_slots.end(), src.triangle_material_slots,
                             src.triangle_material_slots + src.triangle_count);
        }
        scene->meshes.push_back(m);
        scene->mesh_positions.emplace_back(src.positions, src.positions + 3ull * src.vertex_count);
        scene->mesh_tris.emplace_back(src.triangles, src.triangles + 3ull * src.triangle_count);
        scene->mesh_tri_slots.emplace_back();
        if (src.triangle_material_slots)
            scene->mesh_tri_slots.back().assign(src.triangle_material_slots,
                                                src.triangle_material_slots + src.triangle_count);
        scene->mesh_slots.emplace_back(src.material_slots, src.material_slots + src.material_slot_count);
        mesh_prim_offset.push_back(prim_offset);
        prim_offset += src.triangle_count;
    }
    scene->mesh_prim_offsets = mesh_prim_offset;
    // BLASes and instances (build_cycles_instance_intersection_plan,
    // src/luisa/path_tracer_scene_geometry.cpp:272-318)
    std::vector<Node4> nodes(1);  // node 0 reserved for the TLAS root
    std::vector<LeafTri> tris;
    std::vector<Instance> instances;
    std::vector<uint32_t> overrides;
    std::vector<uint32_t> shared_root(d->mesh_count, kNone);
    std::vector<std::array<float, 6>> shared_bounds(d->mesh_count);
    std::vector<uint32_t> shared_depth(d->mesh_count, 0u);
    std::vector<BuildItem> tlas_items;
    std::vector<uint32_t> tlas_instance;
    uint32_t max_blas_depth = 0;
    auto build_blas = [&](const float *pos, const uint32_t *tri, uint32_t tcount, uint32_t &root, float lo[3],
                          float hi[3], uint32_t &depth) {
        std::vector<BuildItem> items(tcount);
        for (uint32_t t = 0; t < tcount; ++t) {
            for (int a = 0; a < 3; ++a) {
                const float x = pos[3u * tri[3u * t] + a], y = pos[3u * tri[3u * t + 1] + a],
                            z = pos[3u * tri[3u * t + 2] + a];
                items[t].lo[a] = std::min(x, std::min(y, z));
                items[t].hi[a] = std::max(x, std::max(y, z));
            }
        }
        const uint32_t base = static_cast<uint32_t>(nodes.size());
        const uint32_t tri_base = static_cast<uint32_t>(tris.size());
        Bvh4 b = build_bvh4(items, base, 2u, bvh_width(backend->options));
        for (auto &n : b.nodes)
            for (int i = 0; i < 4; ++i)
                if (n.count[i] != kEmptySlot && n.count[i] > 0u) n.child[i] += tri_base;
        for (uint32_t idx : b.order) {
            LeafTri lt{};
            const uint32_t *t3 = tri + 3u * idx;
            lt.p0x = pos[3u * t3[0]];
            lt.p0y = pos[3u * t3[0] + 1];
            lt.p0z = pos[3u * t3[0] + 2];
            std::memcpy(&lt.prim_bits, &idx, 4);
            lt.p1x = pos[3u * t3[1]];
            lt.p1y = pos[3u * t3[1] + 1];
            lt.p1z = pos[3u * t3[1] + 2];
            lt.p2x = pos[3u * t3[2]];
            lt.p2y = pos[3u * t3[2] + 1];
            lt.p2z = pos[3u * t3[2] + 2];
            tris.push_back(lt);
        }
        nodes.insert(nodes.end(), b.nodes.begin(), b.nodes.end());
        root = base;
        std::memcpy(lo, b.lo, sizeof(float) * 3);
        std::memcpy(hi, b.hi, sizeof(float) * 3);
        depth = b.depth;
    };
    scene->overrides.resize(d->instance_count);
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const psx_instance &src = d->instances[i];
        scene->instances.push_back(src);
        const uint32_t mi = src.mesh;
        for (uint32_t o = 0; o < src.material_override_count; ++o)
            if (src.material_overrides[o] >= d->material_count)
                return fail("instance " + std::to_string(i) + " overrides with a missing material");
        scene->overrides[i].assign(src.material_overrides, src.material_overrides + src.material_override_count);
        Instance g{};
        float w2o[16], n2w[16];
        cycles_inverse(src.transform, w2o);
        transpose4(w2o, n2w);
        g.w2o = to_affine(w2o);
        g.o2w = to_affine(src.transform);
        g.n2w = to_affine(n2w);
        g.mesh = mi;
        g.visibility = src.visibility_mask;
        g.prim_offset = mesh_prim_offset[mi];
        g.override_offset = static_cast<uint32_t>(overrides.size());
        g.override_count = src.material_override_count;
        if (!(src.shadow_terminator_geometry_offset >= 0.0f) || !std::isfinite(src.shadow_terminator_geometry_offset))
            return fail("instance " + std::to_string(i) + " has an invalid shadow terminator geometry offset");
        g.terminator_offset = src.shadow_terminator_geometry_offset;
        g.shade_class = instance_shade_class(src, scene->meshes[mi], slots, material_class);
        overrides.insert(overrides.end(), src.material_overrides, src.material_overrides + src.material_override_count);
        const Mesh &mesh = scene->meshes[mi];
        float lo[3], hi[3];
        uint32_t depth = 0;
        g.applied = users[mi] == 1u ? 1u : 0u;
        if (g.applied) {
            // single-user mesh: host-transformed final world support positions
            const auto &op = scene->mesh_positions[mi];
            std::vector<float> wp(op.size());
            for (size_t v = 0; v < op.size() / 3u; ++v) {
                const Vec3 p = transform_point(src.transform, v3(op[3 * v], op[3 * v + 1], op[3 * v + 2]));
                wp[3 * v] = p.x;
                wp[3 * v + 1] = p.y;
                wp[3 * v + 2] = p.z;
            }
            g.ipos_offset = static_cast<uint32_t>(positions.size() / 3u);
            positions.insert(positions.end(), wp.begin(), wp.end());
            build_blas(wp.data(), scene->mesh_tris[mi].data(), mesh.triangle_count, g.blas_root, lo, hi, depth);
        } else {
            g.ipos_offset = mesh.vertex_offset;
            if (shared_root[mi] == kNone) {
                float slo[3], shi[3];
                build_blas(scene->mesh_positions[mi].data(), scene->mesh_tris[mi].data(), mesh.triangle_count,
                           shared_root[mi], slo, shi, shared_depth[mi]);
                std::copy(slo, slo + 3, shared_bounds[mi].begin());
                std::copy(shi, shi + 3, shared_bounds[mi].begin() + 3);
            }
            g.blas_root = shared_root[mi];
            depth = shared_depth[mi];
            // world box of the object-space root box
            float wl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, wh[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (int c = 0; c < 8; ++c) {
                const Vec3 corner = v3((c & 1) ? shared_bounds[mi][3] : shared_bounds[mi][0],
                                       (c & 2) ? shared_bounds[mi][4] : shared_bounds[mi][1],
                                       (c & 4) ? shared_bounds[mi][5] : shared_bounds[mi][2]);
                const Vec3 w = transform_point(src.transform, corner);
                wl[0] = std::min(wl[0], w.x);
                wl[1] = std::min(wl[1], w.y);
                wl[2] = std::min(wl[2], w.z);
                wh[0] = std::max(wh[0], w.x);
                wh[1] = std::max(wh[1], w.y);
                wh[2] = std::max(wh[2], w.z);
            }
            for (int a = 0; a < 3; ++a) {
                const float s = std::max(std::max(std::fabs(wl[a]), std::fabs(wh[a])), wh[a] - wl[a]);
                lo[a] = wl[a] - (s * 1.0e-5f + 1.0e-30f);
                hi[a] = wh[a] + (s * 1.0e-5f + 1.0e-30f);
            }
        }
        max_blas_depth = std::max(max_blas_depth, depth);
        if (mesh.triangle_count > 0u && lo[0] <= hi[0]) {
            BuildItem it;
            std::copy(lo, lo + 3, it.lo);
            std::copy(hi, hi + 3, it.hi);
            tlas_items.push_back(it);
            tlas_instance.push_back(i);
        }
        instances.push_back(g);
        scene->triangles += mesh.triangle_count;
    }
    // TLAS, one instance per leaf item
    uint32_t tlas_depth = 0;
    std::vector<uint32_t> tlas_leaf_items;
    bool has_tlas = !tlas_items.empty();
    uint32_t tlas_begin = static_cast<uint32_t>(nodes.size());  // node 0 and [tlas_begin, end) are TLAS nodes
    if (has_tlas) {
        const uint32_t base = static_cast<uint32_t>(nodes.size());
        Bvh4 t = build_bvh4(tlas_items, base, 1u, bvh_width(backend->options));
        // the TLAS root must live at node 0: place it there and append the rest
        for (uint32_t idx : t.order) tlas_leaf_items.push_back(tlas_instance[idx]);
        // inner child indices point at base + local index; local 0 moves to slot 0
        std::vector<Node4> tn = t.nodes;
        for (auto &n : tn)
            for (int i = 0; i < 4; ++i) {
                if (n.count[i] == 0u) n.child[i] -= 1u;  // local k >= 1 lives at base + k - 1
                else if (n.count[i] != kEmptySlot) n.child[i] = tlas_leaf_items[n.child[i]];  // one instance
            }
        nodes[0] = tn[0];
        nodes.insert(nodes.end(), tn.begin() + 1, tn.end());
        tlas_depth = t.depth;
    }
    scene->tlas_depth = tlas_depth;
    scene->blas_depth = max_blas_depth;
    // stack bound: a TLAS node of w children pushes at most w entries
    // (instance markers and inner children), a BLAS node at most w - 1
    const uint32_t width = bvh_width(backend->options);
    scene->max_stack = width * tlas_depth + (width - 1u) * max_blas_depth + 2u;
    if (tris.size() >= kMaxLeafTriangles || nodes.size() >= (1ull << 30) || d->instance_count >= (1u << 30))
        return fail("scene exceeds the traversal entry encoding (" + std::to_string(tris.size()) +
                    " leaf triangles, " + std::to_string(nodes.size()) + " nodes)");
    if (scene->max_stack > kTraversalStackDepth)
        return fail("BVH too deep for the traversal stack (" + std::to_string(scene->max_stack) + " > " +
                    std::to_string(kTraversalStackDepth) + ")");
    // upload
    Scene &s = *scene;
    s.nodes.upload(quantize_nodes(nodes, [&](uint32_t i) { return has_tlas && (i == 0u || i >= tlas_begin); }));
    tris.push_back(LeafTri{});  // padding: a leaf step fetches 64 bytes from its 48-byte triangle
    s.tris.upload(tris);
    s.instances_buf.upload(instances);
    s.meshes_buf.upload(scene->meshes);
    s.positions.upload(positions);
    s.indices.upload(indices);
    s.normals.upload(normals);
    s.smooth.upload(smooth);
    s.slots.upload(slots);
    s.tri_slots.upload(tri_slots);
    s.overrides_buf.upload(overrides);
    s.leaves.upload(leaves);
    hip_check(hipStreamCreateWithFlags(&s.table_stream, hipStreamNonBlocking));
    hip_check(hipStreamCreateWithFlags(&s.query_stream, hipStreamNonBlocking));
    s.tables = std::make_shared<ParamTables>(s.table_stream);
    s.upload_materials(*s.tables);
    SceneView &v = s.view;
    v.nodes = s.nodes.as<QNode4>();
    v.tris = s.tris.as<LeafTri>();
    v.instances = s.instances_buf.as<Instance>();
    v.meshes = s.meshes_buf.as<Mesh>();
    v.positions = s.positions.as<float>();
    v.indices = s.indices.as<uint32_t>();
    v.normals = s.normals.as<float>();
    v.smooth = s.smooth.as<uint8_t>();
    v.slots = s.slots.as<uint32_t>();
    v.tri_slots = s.tri_slots.as<uint32_t>();
    v.overrides = s.overrides_buf.as<uint32_t>();
    v.leaves = s.leaves.as<Leaf>();
    v.tlas_root = 0u;
    v.has_tlas = has_tlas ? 1u : 0u;
    v.instance_count = d->instance_count;
    v.material_count = d->material_count;
    // background lights fold into the constant world radiance
    // (src/luisa/path_tracer_scene.cpp:1723-1734)
    float bg[3] = {d->world_color[0], d->world_color[1], d->world_color[2]};
    for (uint32_t i = 0; i < d->light_count; ++i) {
        const psx_light &l = d->lights[i];
        if (l.type == PSX_LIGHT_POINT || l.type == PSX_LIGHT_SPOT || l.type == PSX_LIGHT_AREA ||
            l.type == PSX_LIGHT_DISTANT) {
            s.lamp_list.push_back(make_lamp(l));
            continue;
        }
        if (l.type != PSX_LIGHT_BACKGROUND) return fail("light " + std::to_string(i) + ": unknown light type");
        bg[0] += l.color[0] * l.power;
        bg[1] += l.color[1] * l.power;
        bg[2] += l.color[2] * l.power;
    }
    v.background[0] = bg[0];
    v.background[1] = bg[1];
    v.background[2] = bg[2];
    v.world_visibility = d->world_visibility_mask;
    s.lamps.upload(s.lamp_list);
    v.lamps = s.lamps.as<Lamp>();
    {
        std::vector<float4> spheres;
        std::vector<uint32_t> bits;
        for (const Lamp &l : s.lamp_list) {
            // an area lamp lies in the ball of its half diagonal about the centre
            const float r = l.type == kLampArea ? 0.5f * std::sqrt(l.size_u * l.size_u + l.size_v * l.size_v)
                                                : l.radius;
            spheres.push_back(make_float4(l.position[0], l.position[1], l.position[2], r));
            const bool intersectable = l.type == kLampArea ? l.size_u * l.size_v != 0.0f
                                                           : l.type != kLampDistant && l.radius > 0.0f;
            bits.push_back((l.visibility & 0xffu) | (intersectable ? kLampBitIntersectable : 0u) |
                           ((l.flags & kLampUseMis) != 0u ? kLampBitMis : 0u));
        }
        s.lamp_spheres.upload(spheres);
        s.lamp_bits.upload(bits);
        v.lamp_sphere = s.lamp_spheres.as<float4>();
        v.lamp_bits = s.lamp_bits.as<uint32_t>();
        v.lamp_nodes = nullptr;
        v.lamp_order = nullptr;
        v.lamp_node_count = 0u;
        if (spheres.size() >= kLampBvhMinLamps) {
            std::vector<LampNode> lnodes;
            std::vector<uint32_t> order;
            build_lamp_bvh(spheres, lnodes, order);
            s.lamp_nodes.upload(lnodes);
            s.lamp_order.upload(order);
            v.lamp_nodes = s.lamp_nodes.as<LampNode>();
            v.lamp_order = s.lamp_order.as<uint32_t>();
            v.lamp_node_count = static_cast<uint32_t>(lnodes.size());
        }
    }
    v.lamp_count = static_cast<uint32_t>(s.lamp_list.size());
    s.rebuild_emitters(*s.tables);
    s.refresh_material_tables(*s.tables);
    hip_check(hipEventSynchronize(s.tables->ready));
    *out = scene.release();
    return PSX_OK;
}

// While a stream capture in global mode runs anywhere in the process, HIP
// refuses allocations, frees, event queries and synchronous copies from
// every thread whose own mode is global (the default).  The library's
// captures use thread-local mode, and its calls only need to stay out of
// captures on their own thread: this scope switches the calling thread to
// thread-local mode and restores the caller's mode on exit.
class ThreadLocalCaptureMode {
  public:
    ThreadLocalCaptureMode() { ok_ = hipThreadExchangeStreamCaptureMode(&saved_) == hipSuccess; }
    ~ThreadLocalCaptureMode() {
        if (ok_) (void)hipThreadExchangeStreamCaptureMode(&saved_);
    }
    ThreadLocalCaptureMode(const ThreadLocalCaptureMode &) = delete;
    ThreadLocalCaptureMode &operator=(const ThreadLocalCaptureMode &) = delete;

  private:
    hipStreamCaptureMode saved_ = hipStreamCaptureModeThreadLocal;
    bool ok_ = false;
};

template <typename F>
int guarded(F &&f) {
    ThreadLocalCaptureMode mode;
    try {
        return f();
    } catch (const DeviceError &e) {
        return e.status;
    } catch (const std::bad_alloc &) {
        return PSX_ERROR_OUT_OF_MEMORY;
    } catch (...) {
        return PSX_ERROR_DEVICE;
    }
}

}  // namespace

// ===========================================================================
// render server: continuous batching of requests over resident scenes
// ===========================================================================
hipError_t launch_batch_begin(const BatchItem *items, uint32_t n_items, uint32_t paths, uint32_t iterations,
                              const WavefrontView &w, hipStream_t stream);
hipError_t launch_batch_iteration(const SceneView &s, const KernelParams &k, const float4 *sobol,
                                  const WavefrontView &w, uint32_t it, Counters *counters, uint32_t *spill,
                                  float share, const GeneratedShade *generated, hipStream_t stream);
hipError_t launch_batch_end(const BatchItem *items, uint32_t n_items, uint32_t pixels, const WavefrontView &w,
                            hipStream_t stream);

namespace {

using ServerClock = std::chrono::steady_clock;

// Requests can share a batch when they render the same scene version with
// the same integrator fields (everything the trace and shade kernels read).
struct BatchKey {
    const Scene *scene = nullptr;
    const ParamTables *tables = nullptr;
    uint32_t u[13] = {};
    float f[4] = {};
    bool operator==(const BatchKey &o) const {
        return scene == o.scene && tables == o.tables && std::memcmp(u, o.u, sizeof(u)) == 0 &&
               std::memcmp(f, o.f, sizeof(f)) == 0;
    }
};

BatchKey batch_key(const Scene *scene, const ParamTables *tables, const KernelParams &k) {
    BatchKey b;
    b.scene = scene;
    b.tables = tables;
    const uint32_t u[13] = {k.max_bounces,      k.min_bounces,    k.max_diffuse,     k.max_glossy, k.max_transmission,
                            k.transparent_min,  k.transparent_max, k.max_path_steps, k.transparent_background,
                            k.nee,              k.mis_mode,        k.use_tree,        k.sequence_size};
    std::memcpy(b.u, u, sizeof(u));
    const float f[4] = {k.clamp_direct, k.clamp_indirect, k.light_inv_rr_threshold, k.pass_alpha_threshold};
    std::memcpy(b.f, f, sizeof(f));
    return b;
}

struct ServerRequest {
    uint64_t id = 0;
    Scene *scene = nullptr;
    std::shared_ptr<ParamTables> tables;  // parameter version bound when the worker took the request
    SceneView view{};
    Session film;                         // film, settings, kernel parameters, filter table
    uint32_t first = 0, count = 0, total = 0;
    uint32_t next = 0;                    // samples handed to batches
    uint32_t done = 0;                    // samples accumulated
    uint32_t padded = 0;
    uint32_t batches = 0;
    bool in_flight = false, viewport = false, admitted = false;
    std::atomic<bool> cancel{false};
    uint32_t state = PSX_REQUEST_QUEUED;
    uint64_t version = 0;     // the scene version this request renders
    bool bound = false;       // version tables and kernel parameters known
    BatchKey key;
    size_t film_bytes = 0;
    hipEvent_t film_ready = nullptr;  // the cleared film, on the control stream
    const float *filter = nullptr;    // the server's table for the request's pixel filter
    int error = PSX_OK;               // why a failed request failed
    std::string message;

    ~ServerRequest() {
        if (film_ready) (void)hipEventDestroy(film_ready);
    }
    double submit_ms = 0.0, admit_ms = -1.0, first_launch_ms = -1.0, done_ms = -1.0;
};

struct ServerOp {
    std::shared_ptr<ServerRequest> request;  // a submission, or
    Scene *scene = nullptr;                  // an edit of `material` on `scene`
    uint32_t material = 0;
    std::vector<float> params;
    uint64_t version = 0;                    // the edit's version number on its scene
};

// Parameter versions of one scene as the server sees them: version n is the
// scene after the first n edits taken by the worker.  The edit thread builds
// them in order; a request binds the version its submission position names.
struct SceneVersions {
    uint64_t requested = 0;   // edits taken
    uint64_t published = 0;   // edits applied and published
    std::map<uint64_t, std::pair<std::shared_ptr<ParamTables>, SceneView>> history;
};

// One stream with its own path state: a batch in flight at a time.
// Iterations of a batch queued on its stream ahead of the device.
constexpr uint32_t kIterationsAhead = 3u;
// fraction of the resident blocks a final-frame batch launches with while
// viewport work waits or runs
constexpr float kViewportShare = 0.25f;
// final-frame batches fill their arena once no viewport request has been
// taken, waited or run for this long
constexpr double kViewportQuietMs = 1000.0;
// edit threads: parameter versions of this many scenes build at once
constexpr uint32_t kEditThreads = 3u;
constexpr uint32_t kMaxIterations = 1025u;

struct Arena {
    WavefrontState state;
    DeviceBuffer items, spill, counters;
    BatchItem *host_items = nullptr;  // pinned staging for the item table
    Counters *host_counters = nullptr;
    WavefrontCounters *host_queues = nullptr;  // pinned: queue sizes after each iteration
    uint32_t item_capacity = 0;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, done = nullptr;
    hipEvent_t ring[kIterationsAhead + 1u] = {};
    GeneratedShade generated{};   // the batch scene's generated shade kernel, if the backend uses them
    bool busy = false;
    bool ending = false;          // accumulation queued; `done` ends the batch
    bool held = false;            // paused for viewport work at least once
    uint32_t iterations = 0;      // at most k.max_path_steps + 1
    uint32_t next_it = 0;         // next iteration to queue
    uint32_t completed_it = 0;    // iterations known to be finished
    // what the iterations and the end of the batch launch with
    SceneView view{};
    KernelParams k{};
    const float4 *sobol = nullptr;
    uint32_t n_items = 0, pixels = 0;
    bool viewport_only = false;
    bool viewport_batch = false;  // the batch in flight serves viewport requests
    BatchKey key;
    std::vector<std::pair<std::shared_ptr<ServerRequest>, uint32_t>> members;
    std::shared_ptr<ParamTables> hold;
    uint64_t paths = 0;
    size_t bytes = 0;

    ~Arena() {
        for (hipEvent_t e : ring)
            if (e) (void)hipEventDestroy(e);
        if (host_queues) (void)hipHostFree(host_queues);
        if (host_items) (void)hipHostFree(host_items);
        if (host_counters) (void)hipHostFree(host_counters);
        if (start) (void)hipEventDestroy(start);
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace

struct Server {
    Backend *backend = nullptr;
    psx_server_options options{};
    std::vector<std::unique_ptr<Arena>> arenas;
    std::map<uint32_t, std::unique_ptr<DeviceBuffer>> sobol;  // by sequence size
    std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<DeviceBuffer>> filters;  // by (filter, width bits)
    hipStream_t control = nullptr;                            // film clears, edits

    std::mutex mutex;
    std::condition_variable wake, finished;
    std::vector<ServerOp> ops;                                // submission order
    std::map<uint64_t, std::shared_ptr<ServerRequest>> requests;
    std::vector<std::shared_ptr<ServerRequest>> active;       // taken, not finished (worker only)
    uint64_t next_id = 1;
    bool stopping = false;
    std::thread worker;
    int worker_error = PSX_OK;
    std::string worker_message;  // the failing call, as psx_last_error reports it
    // edits run on their own threads so that host-side rebuilds and table
    // uploads never delay batch feeding; edits of different scenes build at
    // once, those of one scene in submission order
    std::vector<std::thread> editors;
    std::deque<ServerOp> edits;
    std::set<const Scene *> editing;  // scenes with an edit being built
    std::condition_variable edit_wake;
    std::map<Scene *, SceneVersions> versions;  // guarded by `mutex`

    // device time per path of recent batches, per key: final-frame batches
    // are sized so that one takes about options.target_batch_ms
    std::vector<std::pair<BatchKey, double>> cost_ns;

    ServerClock::time_point epoch = ServerClock::now();
    double last_viewport_ms = -1.0e30;  // viewport work last taken, waiting or running (worker only)
    psx_serving_stats stats{};
    size_t fixed_bytes = 0, film_bytes = 0;
    ServerClock::time_point busy_since{};
    uint32_t busy_arenas = 0;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(ServerClock::now() - epoch).count();
    }

    ~Server() {
        {
            // running batches finish; everything else is cancelled
            std::lock_guard<std::mutex> lock(mutex);
            stopping = true;
            for (auto &entry : requests) entry.second->cancel.store(true);
        }
        wake.notify_all();
        edit_wake.notify_all();
        if (worker.joinable()) worker.join();
        for (std::thread &t : editors)
            if (t.joinable()) t.join();
        // films are released on the control stream: drop them before it goes
        ops.clear();
        active.clear();
        requests.clear();
        arenas.clear();
        if (control) {
            (void)hipStreamSynchronize(control);
            (void)hipStreamDestroy(control);
        }
    }

    // HBM the server's work keeps resident besides arenas and films: the
    // geometry of every scene it renders and each distinct parameter version
    // (and emitter table) still reachable from the scene, the version history
    // or a bound request.  Caller holds `mutex`.
    size_t resident_bytes() const {
        std::set<const void *> seen;
        size_t n = 0;
        auto add = [&](const std::shared_ptr<ParamTables> &t) {
            if (!t || !seen.insert(t.get()).second) return;
            n += t->bytes();
            if (t->light && seen.insert(t->light.get()).second) n += t->light->bytes();
        };
        for (const auto &sv : versions) {
            n += sv.first->geometry_bytes();
            add(sv.first->tables);
            for (const auto &h : sv.second.history) add(h.second.first);
        }
        for (const auto &rp : active) add(rp->tables);
        return n;
    }

    double &cost_of(const BatchKey &key) {
        for (auto &c : cost_ns)
            if (c.first == key) return c.second;
        cost_ns.emplace_back(key, 5.0);  // first guess, refined by every batch
        return cost_ns.back().second;
    }

    const float *filter_table_for(uint32_t filter, float width) {
        uint32_t bits;
        std::memcpy(&bits, &width, sizeof(bits));
        const auto key = std::make_pair(filter, bits);
        auto it = filters.find(key);
        if (it == filters.end()) {
            auto buf = std::make_unique<DeviceBuffer>();
            buf->upload(pixel_filter_table(filter, width));
            it = filters.emplace(key, std::move(buf)).first;
        }
        return it->second->as<float>();
    }

    const float4 *sobol_table_for(uint32_t size) {
        auto it = sobol.find(size);
        if (it == sobol.end()) {
            auto buf = std::make_unique<DeviceBuffer>();
            buf->upload(sobol_table(size));
            it = sobol.emplace(size, std::move(buf)).first;
        }
        return it->second->as<float4>();
    }

    // -- worker ---------------------------------------------------------------
    // caller holds `mutex`
    SceneVersions &versions_of(Scene *scene) {
        auto it = versions.find(scene);
        if (it == versions.end()) {
            // first use: version 0 is the scene as it is now (no edit of
            // this server can be running on it yet)
            it = versions.emplace(scene, SceneVersions{}).first;
            it->second.history[0] = std::make_pair(scene->tables, scene->view);
        }
        return it->second;
    }

    void take(ServerOp &op) {
        std::lock_guard<std::mutex> lock(mutex);
        if (op.request) {
            ServerRequest &r = *op.request;
            r.version = versions_of(r.scene).requested;
            r.film.scene = r.scene;
            const psx_render_settings &rs = r.film.settings;
            r.film.window_w = rs.window_width ? rs.window_width : rs.width;
            r.film.window_h = rs.window_height ? rs.window_height : rs.height;
            r.padded = ((r.film.window_w + 7u) >> 3) * ((r.film.window_h + 7u) >> 3) * 64u;
            const size_t pixels = static_cast<size_t>(r.film.window_w) * r.film.window_h;
            r.viewport = pixels <= options.viewport_max_pixels && r.count <= options.viewport_max_samples;
            if (r.viewport) last_viewport_ms = now_ms();
            r.film_bytes = pixels * 52u + 4096u;
            active.push_back(op.request);
            return;
        }
        // parameter-only edit: version n + 1 of its scene, for the requests
        // taken from now on; built by the edit thread
        op.version = ++versions_of(op.scene).requested;
        edits.push_back(std::move(op));
        edit_wake.notify_all();
    }

    // binds requests whose scene version is published (worker only)
    bool bind() {
        bool progress = false;
        std::lock_guard<std::mutex> lock(mutex);
        for (auto &rp : active) {
            ServerRequest &r = *rp;
            if (r.bound) continue;
            SceneVersions &v = versions_of(r.scene);
            if (v.published < r.version) continue;
            const auto &entry = v.history.at(r.version);
            r.tables = entry.first;
            r.view = entry.second;
            KernelParams k = kernel_params(*r.scene, r.view, r.film.settings);
            k.sequence_size = sequence_size_for(r.total);
            r.film.params = k;
            r.key = batch_key(r.scene, r.tables.get(), k);
            r.bound = true;
            progress = true;
        }
        // versions no request can still bind
        for (auto &sv : versions) {
            uint64_t keep = sv.second.requested;
            for (auto &rp : active)
                if (!rp->bound && rp->scene == sv.first) keep = std::min(keep, rp->version);
            keep = std::min(keep, sv.second.published);
            auto &h = sv.second.history;
            while (!h.empty() && h.begin()->first < keep) h.erase(h.begin());
        }
        return progress;
    }

    void edit_loop() {
        ThreadLocalCaptureMode mode;
        try {
            hip_check(hipSetDevice(backend->device));
            for (;;) {
                ServerOp op;
                {
                    // the oldest edit whose scene has no earlier edit
                    // pending or being built
                    std::unique_lock<std::mutex> lock(mutex);
                    auto ready = [&] {
                        std::set<const Scene *> earlier;
                        for (auto it = edits.begin(); it != edits.end(); ++it) {
                            if (!editing.count(it->scene) && !earlier.count(it->scene)) return it;
                            earlier.insert(it->scene);
                        }
                        return edits.end();
                    };
                    // a closing server cancels every request, so pending
                    // edits are dropped
                    edit_wake.wait(lock, [&] { return stopping || ready() != edits.end(); });
                    if (stopping) return;
                    const auto it = ready();
                    op = std::move(*it);
                    edits.erase(it);
                    editing.insert(op.scene);
                }
                HostMaterial &m = op.scene->materials[op.material];
                m.params = op.params;
                const bool emitted = m.may_emit;
                refresh_may_emit(m);
                op.scene->gpu_materials[op.material].may_emit = m.may_emit ? 1u : 0u;
                // stream-ordered allocations and copies on the scene's
                // table stream, which never synchronises with the null
                // stream, so edits need not wait for the capture gate
                op.scene->apply_parameter_version(emitted || m.may_emit);
                {
                    std::lock_guard<std::mutex> lock(mutex);
                    SceneVersions &v = versions_of(op.scene);
                    v.history[op.version] = std::make_pair(op.scene->tables, op.scene->view);
                    v.published = op.version;
                    stats.edits += 1u;
                    editing.erase(op.scene);
                }
                wake.notify_all();
                edit_wake.notify_all();
            }
        } catch (const DeviceError &e) {
            std::lock_guard<std::mutex> lock(mutex);
            worker_error = e.status;
            worker_message = "server edit thread: " + g_last_error;
        } catch (...) {
            std::lock_guard<std::mutex> lock(mutex);
            worker_error = PSX_ERROR_DEVICE;
            worker_message = "server edit thread: unexpected exception";
        }
        wake.notify_all();
        edit_wake.notify_all();
    }

    void finish(ServerRequest &r, uint32_t state) {
        std::lock_guard<std::mutex> lock(mutex);
        r.state = state;
        r.done_ms = now_ms();
        if (r.admitted) film_bytes -= r.film_bytes;
        if (state != PSX_REQUEST_DONE) {
            // a cancelled or failed request keeps no film
            r.film.combined.reset();
            r.film.normal.reset();
            r.film.albedo.reset();
            r.film.sample_count.reset();
            r.film.filter.reset();
        }
        if (state == PSX_REQUEST_DONE) stats.completed += 1u;
        finished.notify_all();
    }

    bool retire() {
        bool progress = false;
        for (auto &a : arenas) {
            if (!a->busy || !a->ending) continue;
            const hipError_t q = hipEventQuery(a->done);
            if (q == hipErrorNotReady) continue;
            hip_check(q);
            progress = true;
            a->busy = false;
            a->ending = false;
            busy_arenas -= 1u;
            if (busy_arenas == 0u)
                stats.busy_ms +=
                    std::chrono::duration<double, std::milli>(ServerClock::now() - busy_since).count();
            stats.rays += a->host_counters->rays;
            float elapsed = 0.0f;
            if (a->paths > 0u && !a->held && hipEventElapsedTime(&elapsed, a->start, a->done) == hipSuccess) {
                // concurrent batches inflate this, which only errs towards
                // shorter batches
                double &c = cost_of(a->key);
                c = 0.7 * c + 0.3 * (static_cast<double>(elapsed) * 1.0e6 / static_cast<double>(a->paths));
            }
            for (auto &[r, n] : a->members) {
                {
                    std::lock_guard<std::mutex> lock(mutex);
                    r->in_flight = false;
                    r->done += n;
                    r->film.rendered = r->done;
                }
                if (r->done == r->count) finish(*r, PSX_REQUEST_DONE);
            }
            a->members.clear();
            a->hold.reset();
        }
        // cancelled requests leave as soon as no chunk of theirs is running
        for (auto &r : active)
            if (r->cancel.load() && !r->in_flight && r->state < PSX_REQUEST_DONE) finish(*r, PSX_REQUEST_CANCELLED);
        const size_t before = active.size();
        {
            // the worker is the only writer of `active`; other threads read
            // it under the mutex (resident_bytes)
            std::lock_guard<std::mutex> lock(mutex);
            active.erase(std::remove_if(active.begin(), active.end(),
                                        [](const std::shared_ptr<ServerRequest> &r) {
                                            return r->state >= PSX_REQUEST_DONE && !r->in_flight;
                                        }),
                         active.end());
        }
        return progress || active.size() != before;
    }

    bool admit() {
        bool progress = false;
        for (auto &rp : active) {
            ServerRequest &r = *rp;
            if (r.admitted || r.state != PSX_REQUEST_QUEUED) continue;
            size_t in_use, films_in_use;
            {
                std::lock_guard<std::mutex> lock(mutex);
                films_in_use = film_bytes;
                in_use = fixed_bytes + film_bytes + resident_bytes();
            }
            // admission in submission order: a request that does not fit
            // holds back the ones after it; one that cannot fit even with no
            // other film resident fails
            if (in_use + r.film_bytes > options.hbm_budget) {
                if (films_in_use == 0u) {
                    r.error = PSX_ERROR_OUT_OF_MEMORY;
                    r.message = "request film of " + std::to_string(r.film_bytes) + " bytes does not fit the HBM budget (" +
                                std::to_string(options.hbm_budget) + " bytes, " + std::to_string(in_use) +
                                " in use with no film resident)";
                    finish(r, PSX_REQUEST_FAILED);
                    progress = true;
                    continue;
                }
                break;
            }
            const size_t pixels = static_cast<size_t>(r.film.window_w) * r.film.window_h;
            // stream-ordered allocation and clearing; the request's first
            // batch waits on film_ready, the worker never blocks here
            for (DeviceBuffer *b : {&r.film.combined, &r.film.normal, &r.film.albedo, &r.film.sample_count})
                b->set_stream(control);
            r.film.combined.allocate(pixels * 16u);
            r.film.normal.allocate(pixels * 16u);
            r.film.albedo.allocate(pixels * 16u);
            r.film.sample_count.allocate(pixels * 4u);
            r.filter = filter_table_for(r.film.settings.pixel_filter, r.film.settings.filter_width);
            hip_check(hipMemsetAsync(r.film.combined.as<void>(), 0, pixels * 16u, control));
            hip_check(hipMemsetAsync(r.film.normal.as<void>(), 0, pixels * 16u, control));
            hip_check(hipMemsetAsync(r.film.albedo.as<void>(), 0, pixels * 16u, control));
            hip_check(hipMemsetAsync(r.film.sample_count.as<void>(), 0, pixels * 4u, control));
            hip_check(hipEventCreateWithFlags(&r.film_ready, hipEventDisableTiming));
            hip_check(hipEventRecord(r.film_ready, control));
            std::lock_guard<std::mutex> lock(mutex);
            r.admitted = true;
            r.state = PSX_REQUEST_RUNNING;
            r.admit_ms = now_ms();
            film_bytes += r.film_bytes;
            stats.hbm_in_use = fixed_bytes + film_bytes + resident_bytes();
            stats.peak_hbm = std::max(stats.peak_hbm, stats.hbm_in_use);
            progress = true;
        }
        return progress;
    }

    static bool runnable(const ServerRequest &r) {
        return r.bound && r.admitted && r.state == PSX_REQUEST_RUNNING && !r.in_flight && !r.cancel.load() &&
               r.next < r.count;
    }

    // Fills `a` with chunks of runnable requests sharing the leader's key:
    // viewport requests first (the whole remainder of each), then the others
    // in submission order, each as many samples as the capacity leaves.
    //
    // While viewport work is waiting or running no new final-frame batch
    // starts.  Final-frame batches are sized from the key's measured cost to
    // last about options.target_batch_ms while viewport work was seen within
    // kViewportQuietMs, and fill the arena otherwise.
    bool launch(Arena &a) {
        ServerRequest *leader = nullptr;
        for (auto &r : active)
            if (runnable(*r) && r->viewport) {
                leader = r.get();
                break;
            }
        const bool viewport = leader != nullptr;
        if (!leader) {
            if (a.viewport_only) return false;
            for (auto &b : arenas)
                if (b->busy && b->viewport_batch) return false;
            for (auto &r : active)
                if (runnable(*r)) {
                    leader = r.get();
                    break;
                }
        }
        if (!leader) return false;
        const BatchKey key = leader->key;
        uint32_t capacity = a.state.view.capacity;
        if (!viewport && now_ms() - last_viewport_ms < kViewportQuietMs) {
            const double target = static_cast<double>(options.target_batch_ms) * 1.0e6 / cost_of(key);
            const uint64_t bounded = std::max<uint64_t>(static_cast<uint64_t>(target), leader->padded);
            capacity = static_cast<uint32_t>(std::min<uint64_t>(bounded, capacity));
        }
        uint32_t paths = 0, pixels = 0, n = 0;
        for (int pass = 0; pass < 2 && n < a.item_capacity; ++pass) {
            const bool want_viewport = pass == 0;
            if (want_viewport != viewport) continue;
            for (auto &rp : active) {
                ServerRequest &r = *rp;
                if (n >= a.item_capacity) break;
                if (!runnable(r) || r.viewport != want_viewport || !(r.key == key)) continue;
                const uint32_t room = (capacity - paths) / r.padded;
                const uint32_t chunk = std::min(r.count - r.next, room);
                if (chunk == 0u) continue;
                BatchItem &it = a.host_items[n];
                it.k = r.film.params;
                it.film = FilmView{r.film.combined.as<float>(), r.film.normal.as<float>(), r.film.albedo.as<float>(),
                                   r.film.sample_count.as<uint32_t>()};
                it.filter = r.filter;
                it.sobol = sobol_table_for(r.film.params.sequence_size);
                it.first = r.first + r.next;
                it.count = chunk;
                it.padded = r.padded;
                it.path_base = paths;
                it.pixel_base = pixels;
                it.pad0 = it.pad1 = it.pad2 = 0u;
                paths += chunk * r.padded;
                pixels += r.padded;
                r.next += chunk;
                r.in_flight = true;
                r.batches += 1u;
                if (r.first_launch_ms < 0.0) r.first_launch_ms = now_ms();
                a.members.emplace_back(rp, chunk);
                ++n;
            }
        }
        if (n == 0u) return false;
        hip_check(hipEventRecord(a.start, a.stream));
        // the parameter version's tables and every member's cleared film
        hip_check(hipStreamWaitEvent(a.stream, leader->tables->ready, 0));
        for (auto &member : a.members) hip_check(hipStreamWaitEvent(a.stream, member.first->film_ready, 0));
        hip_check(hipMemcpyAsync(a.items.as<void>(), a.host_items, n * sizeof(BatchItem), hipMemcpyHostToDevice,
                                 a.stream));
        hip_check(hipMemsetAsync(a.counters.as<void>(), 0, sizeof(Counters), a.stream));
        a.k = a.host_items[0].k;
        a.view = leader->view;
        a.generated = (leader->scene->backend->options.flags & PSX_BACKEND_GENERATED_PROGRAMS) != 0u
                          ? generated_program(*leader->scene, shade_variant(a.view, false))
                          : GeneratedShade{};
        a.sobol = sobol_table_for(a.k.sequence_size);
        a.n_items = n;
        a.pixels = pixels;
        a.iterations = a.k.max_path_steps + 1u;
        a.next_it = 0u;
        a.completed_it = 0u;
        a.ending = false;
        a.held = false;
        hip_check(launch_batch_begin(a.items.as<BatchItem>(), n, paths, a.iterations, a.state.view, a.stream));
        a.hold = leader->tables;
        a.paths = paths;
        a.key = key;
        a.viewport_batch = viewport;
        a.busy = true;
        if (busy_arenas == 0u) busy_since = ServerClock::now();
        busy_arenas += 1u;
        stats.batches += 1u;
        stats.paths += paths;
        feed(a, false);
        return true;
    }

    bool viewport_pending() const {
        for (auto &r : active)
            if (runnable(*r) && r->viewport) return true;
        for (auto &a : arenas)
            if (a->busy && a->viewport_batch && !a->ending) return true;
        return false;
    }

    // Queues the next iterations of a running batch (at most
    // kIterationsAhead ahead of the device), or its end once the queues ran
    // dry or every iteration is queued.  While `hold` (viewport work waits or
    // runs), a final-frame batch's iterations launch on kViewportShare of the
    // device's resident blocks, which leaves compute units free for the
    // viewport batches' kernels.
    bool feed(Arena &a, bool hold) {
        if (!a.busy || a.ending) return false;
        bool progress = false;
        while (a.completed_it < a.next_it) {
            const hipError_t q = hipEventQuery(a.ring[a.completed_it % (kIterationsAhead + 1u)]);
            if (q == hipErrorNotReady) break;
            hip_check(q);
            a.completed_it += 1u;
            progress = true;
        }
        const WavefrontCounters &after = a.host_queues[a.completed_it];
        const bool drained = a.completed_it > 0u && after.trace_count == 0u && after.shade_count == 0u;
        if (drained || a.next_it == a.iterations) {
            // stream order puts the end after every iteration already queued
            hip_check(launch_batch_end(a.items.as<BatchItem>(), a.n_items, a.pixels, a.state.view, a.stream));
            hip_check(hipMemcpyAsync(a.host_counters, a.counters.as<void>(), sizeof(Counters),
                                     hipMemcpyDeviceToHost, a.stream));
            hip_check(hipEventRecord(a.done, a.stream));
            a.ending = true;
            return true;
        }
        const bool shrink = hold && !a.viewport_batch;
        while (a.next_it < a.iterations && a.next_it - a.completed_it < kIterationsAhead) {
            const uint32_t it = a.next_it;
            a.held |= shrink;
            hip_check(launch_batch_iteration(a.view, a.k, a.sobol, a.state.view, it, a.counters.as<Counters>(),
                                             a.spill.as<uint32_t>(), shrink ? kViewportShare : 1.0f,
                                             a.generated.fn ? &a.generated : nullptr, a.stream));
            hip_check(hipMemcpyAsync(&a.host_queues[it + 1u], a.state.view.counters + it + 1u,
                                     sizeof(WavefrontCounters), hipMemcpyDeviceToHost, a.stream));
            hip_check(hipEventRecord(a.ring[it % (kIterationsAhead + 1u)], a.stream));
            a.next_it += 1u;
            progress = true;
        }
        return progress;
    }

    void run() {
        ThreadLocalCaptureMode mode;
        try {
            hip_check(hipSetDevice(backend->device));
            for (;;) {
                std::vector<ServerOp> taken;
                {
                    std::unique_lock<std::mutex> lock(mutex);
                    taken.swap(ops);
                    if (worker_error != PSX_OK) throw DeviceError{worker_error};
                    if (stopping && taken.empty() && active.empty() && busy_arenas == 0u) break;
                }
                bool progress = !taken.empty();
                {
                    std::lock_guard<std::mutex> gate(backend->capture_gate);
                    for (ServerOp &op : taken) take(op);
                    progress |= bind();
                    progress |= retire();
                    progress |= admit();
                    for (auto &a : arenas)
                        if (!a->busy) progress |= launch(*a);
                    const bool hold = viewport_pending();
                    if (hold) last_viewport_ms = now_ms();
                    for (auto &a : arenas) progress |= feed(*a, hold);
                }
                if (stopping && !progress && busy_arenas == 0u) {
                    // nothing left that can run (requests that never fit)
                    for (auto &r : active)
                        if (r->state < PSX_REQUEST_DONE) finish(*r, PSX_REQUEST_CANCELLED);
                    std::lock_guard<std::mutex> lock(mutex);
                    active.clear();
                }
                if (!progress) {
                    std::unique_lock<std::mutex> lock(mutex);
                    wake.wait_for(lock, std::chrono::microseconds(busy_arenas ? 50 : 2000),
                                  [&] { return !ops.empty() || stopping; });
                }
            }
        } catch (const DeviceError &e) {
            std::lock_guard<std::mutex> lock(mutex);
            worker_error = e.status;
            if (worker_message.empty()) worker_message = "server worker: " + g_last_error;
        } catch (...) {
            std::lock_guard<std::mutex> lock(mutex);
            worker_error = PSX_ERROR_DEVICE;
            if (worker_message.empty()) worker_message = "server worker: unexpected exception";
        }
        if (worker_error != PSX_OK) {
            // wait for what is on the device before the films can go
            for (auto &a : arenas)
                if (a->busy) (void)hipEventSynchronize(a->done);
            std::lock_guard<std::mutex> lock(mutex);
            for (auto &r : active)
                if (r->state < PSX_REQUEST_DONE) r->state = PSX_REQUEST_FAILED;
            finished.notify_all();
        }
    }
};

namespace {
}  // namespace
}  // namespace psx

// ===========================================================================
// C ABI
// ===========================================================================
using namespace psx;

struct psx_backend : Backend {};
struct psx_server : Server {};
struct psx_scene : Scene {};
struct psx_session : Session {};

extern "C" {

const char *psx_version(void) { return "psx 0.1 (gfx950)"; }

int psx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void psx_render_settings_default(psx_render_settings *s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->pixel_filter = PSX_FILTER_BOX;
    s->filter_width = 1.0f;
    s->pass_alpha_threshold = 0.5f;
    psx_integrator_settings &i = s->integrator;
    i.max_bounces = 12u;
    i.min_bounces = 0u;
    i.diffuse_bounces = 4u;
    i.glossy_bounces = 4u;
    i.transmission_bounces = 12u;
    i.volume_bounces = 0u;
    i.transparent_min_bounces = 0u;
    i.transparent_max_bounces = 8u;
    i.film_exposure = 1.0f;
    i.light_sampling_threshold = 0.01f;
    i.reflective_caustics = 1u;
    i.refractive_caustics = 1u;
    i.direct_light_sampling = PSX_DIRECT_LIGHT_MIS;
}

int psx_backend_create(int device, const psx_backend_options *options, psx_backend **out) {
    if (!out) return PSX_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return static_cast<int>(PSX_ERROR_NO_DEVICE);
        if (device < 0 || device >= n) return static_cast<int>(PSX_ERROR_INVALID_ARGUMENT);
        hip_check(hipSetDevice(device));
        // address of the allocation-free query scratch on this device
        hip_check(setup_noalloc_scratch());
        auto *b = new psx_backend();
        b->device = device;
        b->options.next_event_estimation = 1u;
        b->options.max_samples_per_dispatch = 4u;
        if (options) {
            b->options = *options;
            const uint32_t widths = PSX_BACKEND_BVH_WIDTH_2 | PSX_BACKEND_BVH_WIDTH_3;
            if (b->options.max_samples_per_dispatch == 0u || (b->options.flags & widths) == widths) {
                delete b;
                return static_cast<int>(PSX_ERROR_INVALID_ARGUMENT);
            }
        }
        *out = b;
        return static_cast<int>(PSX_OK);
    });
}

void psx_backend_destroy(psx_backend *b) {
    ThreadLocalCaptureMode mode;
    delete b;
}

int psx_scene_compile(psx_backend *b, const psx_scene_desc *d, psx_diag_fn diag, void *user, psx_scene **out) {
    if (!b || !d || !out) return PSX_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded([&] {
        hip_check(hipSetDevice(b->device));
        Scene *s = nullptr;
        const int rc = compile(b, d, diag, user, &s);
        if (rc != PSX_OK) return rc;
        // re-home into the ABI type
        *out = reinterpret_cast<psx_scene *>(s);
        return static_cast<int>(PSX_OK);
    });
}

void psx_scene_destroy(psx_scene *s) {
    ThreadLocalCaptureMode mode;
    delete reinterpret_cast<Scene *>(s);
}

int psx_scene_update_parameters(psx_scene *handle, uint32_t material, const float *p, uint32_t count) {
    Scene *s = reinterpret_cast<Scene *>(handle);
    if (!s || !p || material >= s->materials.size()) return PSX_ERROR_INVALID_ARGUMENT;
    HostMaterial &m = s->materials[material];
    if (count != m.params.size()) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->backend->device));
        m.params.assign(p, p + count);
        const bool emitted = m.may_emit;
        refresh_may_emit(m);
        Material &g = s->gpu_materials[material];
        g.may_emit = m.may_emit ? 1u : 0u;
        s->apply_parameter_version(emitted || m.may_emit);
        hip_check(hipEventSynchronize(s->tables->ready));
        return static_cast<int>(PSX_OK);
    });
}

int psx_session_create(psx_scene *handle, const psx_render_settings *rs, psx_session **out) {
    Scene *scene = reinterpret_cast<Scene *>(handle);
    if (!scene || !out) return PSX_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!valid_settings(rs)) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(scene->backend->device));
        auto s = std::make_unique<Session>();
        s->scene = scene;
        hip_check(hipStreamCreate(&s->stream));
        hip_check(hipEventCreate(&s->ev_begin));
        hip_check(hipEventCreate(&s->ev_end));
        // traversal spill columns for every lane the traversal kernels keep resident
        const size_t lanes = std::max(megakernel_lanes(), wavefront_lanes());
        s->spill.allocate(lanes * kTraversalStackDepth * sizeof(uint32_t));
        s->tile_cursor.allocate(sizeof(uint32_t));
        s->count_nodes = (scene->backend->options.flags & PSX_BACKEND_COUNT_TRAVERSAL) != 0u;
        initialize_session(*s, *rs);
        *out = reinterpret_cast<psx_session *>(s.release());
        return static_cast<int>(PSX_OK);
    });
}

void psx_session_destroy(psx_session *s) {
    ThreadLocalCaptureMode mode;
    delete reinterpret_cast<Session *>(s);
}

int psx_session_reset(psx_session *handle, const psx_render_settings *rs) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || !valid_settings(rs)) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        s->cancelled.store(false);
        initialize_session(*s, *rs);
        return static_cast<int>(PSX_OK);
    });
}

namespace {
int check_range(Session *s, uint32_t first, uint32_t count, uint32_t offset, uint32_t total) {
    if (s->cancelled.load()) return PSX_ERROR_CANCELLED;
    if (count == 0u) return PSX_ERROR_RANGE;
    const uint64_t end = static_cast<uint64_t>(first) + offset + count;
    if (total == 0u || end > total || (s->total_aa != 0u && s->total_aa != total)) return PSX_ERROR_RANGE;
    return PSX_OK;
}
}  // namespace

int psx_session_render(psx_session *handle, uint32_t first, uint32_t count, uint32_t offset, uint32_t total,
                       const psx_sink *sink) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        const int rc = check_range(s, first, count, offset, total);
        if (rc != PSX_OK) return rc;
        hip_check(hipSetDevice(s->scene->backend->device));
        s->total_aa = total;
        prepare_sobol(*s, total);
        const int d = dispatch_samples(*s, first + offset, count);
        if (d != PSX_OK) return d;
        if (s->cancelled.load()) return static_cast<int>(PSX_ERROR_CANCELLED);
        return write_passes(*s, sink);
    });
}

void psx_session_cancel(psx_session *handle) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (s) s->cancelled.store(true);
}

int psx_session_stats(const psx_session *handle, psx_render_stats *out) {
    const Session *s = reinterpret_cast<const Session *>(handle);
    if (!s || !out) return PSX_ERROR_INVALID_ARGUMENT;
    *out = s->stats;
    return PSX_OK;
}

int psx_session_render_device(psx_session *handle, uint32_t first, uint32_t count, uint32_t total) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        const int rc = check_range(s, first, count, 0u, total);
        if (rc != PSX_OK) return rc;
        hip_check(hipSetDevice(s->scene->backend->device));
        s->total_aa = total;
        prepare_sobol(*s, total);
        return dispatch_samples(*s, first, count);
    });
}

const char *psx_last_error(void) { return g_last_error.c_str(); }

namespace {

// Closure-leaf layouts of a material array, as scene compilation lowers them.
int material_layouts(const psx_material *materials, uint32_t count, std::vector<LeafLayout> &layouts) {
    if (count && !materials) return PSX_ERROR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < count; ++i) {
        const psx_material &src = materials[i];
        if ((src.instruction_count && !src.instructions) || (src.parameter_count && !src.parameters))
            return PSX_ERROR_INVALID_ARGUMENT;
        std::vector<Leaf> ml;
        std::string error;
        if (!lower_leaves(host_material(src), ml, error)) return PSX_ERROR_INVALID_ARGUMENT;
        (void)program_index(layouts, ml);
    }
    return PSX_OK;
}

void copy_text(const std::string &text, char *out, size_t capacity) {
    if (!out || capacity == 0u) return;
    const size_t n = std::min(capacity - 1u, text.size());
    std::memcpy(out, text.data(), n);
    out[n] = '\0';
}

}  // namespace

int psx_surface_program_source(const psx_material *materials, uint32_t material_count, uint32_t variant,
                               char *source, size_t capacity, size_t *length, uint32_t *programs) {
    if (variant > 3u || !length) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        std::vector<LeafLayout> layouts;
        const int rc = material_layouts(materials, material_count, layouts);
        if (rc != PSX_OK) return rc;
        const std::string text = surface_program_source(layouts, variant);
        *length = text.size();
        if (programs) *programs = static_cast<uint32_t>(layouts.size());
        if (source && capacity <= text.size()) return static_cast<int>(PSX_ERROR_INVALID_ARGUMENT);
        copy_text(text, source, capacity);
        return static_cast<int>(PSX_OK);
    });
}

int psx_surface_program_build(const psx_material *materials, uint32_t material_count, uint32_t variant,
                              const char *arch, char *log, size_t log_capacity, uint64_t *code_bytes) {
    if (variant > 3u) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        std::vector<LeafLayout> layouts;
        const int rc = material_layouts(materials, material_count, layouts);
        if (rc != PSX_OK) return rc;
        const RtcResult r = compile_surface_program(surface_program_source(layouts, variant), arch ? arch : "gfx950");
        copy_text(r.log, log, log_capacity);
        if (code_bytes) *code_bytes = r.code.size();
        return static_cast<int>(r.ok ? PSX_OK : PSX_ERROR_UNSUPPORTED);
    });
}

uint32_t psx_backend_generated_programs(const psx_backend *handle) {
    const Backend *b = reinterpret_cast<const Backend *>(handle);
    return b ? b->programs.compiled() : 0u;
}

void psx_wavefront_tuning_default(psx_wavefront_tuning *t) {
    if (t) *t = psx_wavefront_tuning{8u, 16u, 128u, 256u, 128u, 0u, 1u};
}

int psx_session_set_tuning(psx_session *handle, const psx_wavefront_tuning *t) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || !t || !valid_tuning(*t)) return PSX_ERROR_INVALID_ARGUMENT;
    s->tuning = *t;
    return PSX_OK;
}

int psx_session_get_tuning(const psx_session *handle, psx_wavefront_tuning *t) {
    const Session *s = reinterpret_cast<const Session *>(handle);
    if (!s || !t) return PSX_ERROR_INVALID_ARGUMENT;
    *t = s->tuning;
    return PSX_OK;
}

namespace {

// Device time of one probe dispatch of samples [0, count) into `film`.
float time_probe(Session &s, const FilmView &film, uint32_t count, const GeneratedShade *generated, hipEvent_t a,
                 hipEvent_t b) {
    prepare_wavefront(s, count);
    hip_check(hipEventRecord(a, s.stream));
    hip_check(launch_wavefront(s.scene->view, s.params, s.sobol.as<float4>(), s.filter.as<float>(), film,
                               s.wavefront.view, 0u, count, nullptr, s.counters.as<Counters>(), s.count_nodes,
                               s.spill.as<uint32_t>(), nullptr, generated, s.stream));
    hip_check(hipEventRecord(b, s.stream));
    hip_check(hipEventSynchronize(b));
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, a, b));
    return ms;
}

// One warm-up dispatch, then the faster of two timed ones.
float time_shape(Session &s, const psx_wavefront_tuning &t, const FilmView &film, uint32_t count,
                 const GeneratedShade *generated, hipEvent_t a, hipEvent_t b) {
    s.tuning = t;
    (void)time_probe(s, film, count, generated, a, b);
    return std::min(time_probe(s, film, count, generated, a, b), time_probe(s, film, count, generated, a, b));
}

int autotune(Session &s, uint32_t probe, psx_tuning_report *report) {
    const psx_backend_options &opt = s.scene->backend->options;
    if (opt.integrator != PSX_INTEGRATOR_WAVEFRONT && !s.light_passes) return PSX_ERROR_UNSUPPORTED;
    if (s.total_aa != 0u) {
        probe = std::min(probe, s.total_aa);
    } else {
        prepare_sobol(s, probe);
    }
    // scratch film: the probes never touch the session's
    const size_t pixels = static_cast<size_t>(s.window_w) * s.window_h;
    DeviceBuffer combined, normal, albedo, count, light;
    combined.allocate(pixels * 16u);
    normal.allocate(pixels * 16u);
    albedo.allocate(pixels * 16u);
    count.allocate(pixels * 4u);
    if (s.light_passes) light.allocate(pixels * kLightPasses * 16u);
    const FilmView film{combined.as<float>(), normal.as<float>(), albedo.as<float>(), count.as<uint32_t>(),
                        s.light_passes ? light.as<float>() : nullptr};
    hipEvent_t a = nullptr, b = nullptr;
    hip_check(hipEventCreate(&a));
    hip_check(hipEventCreate(&b));
    psx_tuning_report r{};
    r.probe_samples = probe;
    GeneratedShade generated_shade{};
    const GeneratedShade *generated = nullptr;
    if ((opt.flags & PSX_BACKEND_GENERATED_PROGRAMS) != 0u) {
        generated_shade = generated_program(*s.scene, shade_variant(s.scene->view, s.light_passes));
        generated = &generated_shade;
    }
    try {
        psx_wavefront_tuning best;
        psx_wavefront_tuning_default(&best);
        r.default_ms = time_shape(s, best, film, probe, generated, a, b);
        float best_ms = r.default_ms;
        uint32_t candidates = 1u;
        // one knob at a time; a shape replaces the best only if >= 1 % faster,
        // and the final choice only if it stays >= 1 % faster than the
        // default when both are timed again in turns
        auto consider = [&](psx_wavefront_tuning t) {
            if (std::memcmp(&t, &best, sizeof(t)) == 0) return;
            const float ms = time_shape(s, t, film, probe, generated, a, b);
            ++candidates;
            if (ms < 0.99f * best_ms) {
                best = t;
                best_ms = ms;
            }
        };
        const psx_wavefront_tuning base_tile = best;
        for (uint32_t tw : {16u, 32u, 64u}) {
            psx_wavefront_tuning t = base_tile;
            t.tile_width = tw;
            consider(t);
        }
        const psx_wavefront_tuning base_refill = best;
        for (uint32_t rl : {8u, 24u, 32u, 48u}) {
            psx_wavefront_tuning t = base_refill;
            t.refill_lanes = rl;
            consider(t);
        }
        const psx_wavefront_tuning base_chunk = best;
        for (uint32_t ch : {64u, 256u, 512u}) {
            psx_wavefront_tuning t = base_chunk;
            t.trace_chunk = ch;
            consider(t);
        }
        const psx_wavefront_tuning base_append = best;
        const uint32_t appends[3][2] = {{128u, 64u}, {128u, 128u}, {256u, 256u}};
        for (const auto &qa : appends) {
            psx_wavefront_tuning t = base_append;
            t.queue_chunk = qa[0];
            t.list_chunk = qa[1];
            consider(t);
        }
        {
            psx_wavefront_tuning t = best;
            t.sort_shading = 1u;
            consider(t);
        }
        {
            psx_wavefront_tuning t = best;
            t.xcd_segments = 8u;
            consider(t);
        }
        psx_wavefront_tuning fallback;
        psx_wavefront_tuning_default(&fallback);
        if (std::memcmp(&best, &fallback, sizeof(best)) != 0) {
            // confirm the winner against the default, timed in turns, so a
            // one-off fast probe does not replace the default shape
            float chosen = best_ms, plain = r.default_ms;
            for (int round = 0; round < 3; ++round) {
                plain = std::min(plain, time_shape(s, fallback, film, probe, generated, a, b));
                chosen = std::min(chosen, time_shape(s, best, film, probe, generated, a, b));
            }
            r.default_ms = plain;
            best_ms = chosen;
            if (!(chosen < 0.99f * plain)) {
                best = fallback;
                best_ms = plain;
            }
        }
        r.chosen = best;
        r.chosen_ms = best_ms;
        r.candidates = candidates;
        s.tuning = best;
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (report) *report = r;
    return PSX_OK;
}

}  // namespace

int psx_session_autotune(psx_session *handle, uint32_t probe_samples, psx_tuning_report *report) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || probe_samples == 0u || probe_samples > 64u) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        return autotune(*s, probe_samples, report);
    });
}

int psx_session_read_film(psx_session *handle, float *combined, float *normal, float *albedo, uint32_t *counts) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        const size_t n = static_cast<size_t>(s->window_w) * s->window_h;
        if (combined) hip_check(hipMemcpy(combined, s->combined.as<void>(), n * 16u, hipMemcpyDeviceToHost));
        if (normal) hip_check(hipMemcpy(normal, s->normal.as<void>(), n * 16u, hipMemcpyDeviceToHost));
        if (albedo) hip_check(hipMemcpy(albedo, s->albedo.as<void>(), n * 16u, hipMemcpyDeviceToHost));
        if (counts) hip_check(hipMemcpy(counts, s->sample_count.as<void>(), n * 4u, hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

int psx_session_read_light_passes(psx_session *handle, float *out) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || !out) return PSX_ERROR_INVALID_ARGUMENT;
    if (!s->light_passes) return PSX_ERROR_UNSUPPORTED;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        const size_t n = static_cast<size_t>(s->window_w) * s->window_h;
        hip_check(hipMemcpy(out, s->light.as<void>(), n * kLightPasses * 16u, hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

int psx_session_primary_hits(psx_session *handle, uint32_t sample, uint32_t total, uint32_t *out) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || !out || total == 0u || sample >= total) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        prepare_sobol(*s, total);
        const size_t n = static_cast<size_t>(s->window_w) * s->window_h;
        DeviceBuffer rec;
        rec.allocate(n * 16u);
        hip_check(launch_primary(s->scene->view, s->params, s->sobol.as<float4>(), s->filter.as<float>(), sample,
                                 rec.as<uint32_t>(), s->spill.as<uint32_t>(), s->stream));
        hip_check(hipStreamSynchronize(s->stream));
        hip_check(hipMemcpy(out, rec.as<void>(), n * 16u, hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

int psx_session_trace_events(psx_session *handle, uint32_t sample, uint32_t total, uint32_t *out) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || !out || total == 0u || sample >= total) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        prepare_sobol(*s, total);
        const size_t n = static_cast<size_t>(s->window_w) * s->window_h;
        const size_t words = n * PSX_TRACE_EVENTS * PSX_TRACE_WORDS;
        DeviceBuffer rec;
        rec.allocate(words * 4u);
        hip_check(launch_trace(s->scene->view, s->params, s->sobol.as<float4>(), s->filter.as<float>(), sample,
                               rec.as<uint32_t>(), s->spill.as<uint32_t>(), s->stream));
        hip_check(hipStreamSynchronize(s->stream));
        hip_check(hipMemcpy(out, rec.as<void>(), words * 4u, hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

int psx_session_path_trace(psx_session *handle, uint32_t pixel_x, uint32_t pixel_y, uint32_t sample,
                           uint32_t total, float *out_slots, uint32_t *out_records) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || !out_slots || total == 0u || sample >= total) return PSX_ERROR_INVALID_ARGUMENT;
    // pixel_y counts from the lower row of the full image (Cycles' film convention)
    const KernelParams &k = s->params;
    if (pixel_x >= k.full_width || pixel_y >= k.full_height) return PSX_ERROR_INVALID_ARGUMENT;
    const uint32_t fy = k.full_height - 1u - pixel_y;
    if (pixel_x < k.window_x || fy < k.window_y) return PSX_ERROR_INVALID_ARGUMENT;
    const uint32_t x = pixel_x - k.window_x, y = fy - k.window_y;
    if (x >= s->window_w || y >= s->window_h) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        prepare_sobol(*s, total);
        DeviceBuffer slots, rec;
        slots.allocate(PSX_PATH_TRACE_SLOTS * 16u);
        rec.allocate(PSX_TRACE_EVENTS * PSX_TRACE_WORDS * 4u);
        hip_check(launch_path_trace(s->scene->view, s->params, s->sobol.as<float4>(), s->filter.as<float>(), sample,
                                    x, y, slots.as<float4>(), rec.as<uint32_t>(), s->spill.as<uint32_t>(),
                                    s->stream));
        hip_check(hipStreamSynchronize(s->stream));
        hip_check(hipMemcpy(out_slots, slots.as<void>(), PSX_PATH_TRACE_SLOTS * 16u, hipMemcpyDeviceToHost));
        if (out_records)
            hip_check(hipMemcpy(out_records, rec.as<void>(), PSX_TRACE_EVENTS * PSX_TRACE_WORDS * 4u,
                                hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

int psx_sobol_table(uint32_t total, float *out, uint32_t *out_count) {
    if (total == 0u) return PSX_ERROR_INVALID_ARGUMENT;
    const uint32_t size = sequence_size_for(total);
    const uint32_t count = size * 256u;
    if (out_count) *out_count = count;
    if (out) {
        const std::vector<float> t = sobol_table(size);
        std::memcpy(out, t.data(), t.size() * sizeof(float));
    }
    return PSX_OK;
}

int psx_session_sampler_probe(psx_session *handle, uint32_t sample, uint32_t total, uint32_t first_dim,
                              uint32_t dim_count, uint32_t *out) {
    Session *s = reinterpret_cast<Session *>(handle);
    if (!s || !out || total == 0u || dim_count == 0u || dim_count > 64u) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        hip_check(hipSetDevice(s->scene->backend->device));
        prepare_sobol(*s, total);
        const size_t n = static_cast<size_t>(s->window_w) * s->window_h;
        const size_t words = n * (1u + 4u * dim_count);
        DeviceBuffer rec;
        rec.allocate(words * 4u);
        hip_check(launch_sampler(s->params, s->sobol.as<float4>(), sample, first_dim, dim_count, rec.as<uint32_t>(),
                                 s->stream));
        hip_check(hipStreamSynchronize(s->stream));
        hip_check(hipMemcpy(out, rec.as<void>(), words * 4u, hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

// ---- ray queries and probes -------------------------------------------------
int psx_scene_intersect_device(psx_scene *handle, const psx_ray_query *queries, uint32_t count, psx_ray_hit *hits) {
    Scene *s = reinterpret_cast<Scene *>(handle);
    if (!s || (count && (!queries || !hits))) return PSX_ERROR_INVALID_ARGUMENT;
    if (count == 0u) return PSX_OK;
    return guarded([&] {
        hip_check(hipSetDevice(s->backend->device));
        const uint32_t lanes = query_lanes(count, query_resident_lanes());
        DeviceBuffer cursor, spill;
        cursor.allocate(sizeof(uint32_t));
        spill.allocate(static_cast<size_t>(lanes) * kTraversalStackDepth * sizeof(uint32_t));
        hip_check(launch_intersect(s->view, queries, count, hits, cursor.as<uint32_t>(), spill.as<uint32_t>(), lanes,
                                   s->query_stream));
        hip_check(hipStreamSynchronize(s->query_stream));
        return static_cast<int>(PSX_OK);
    });
}

int psx_scene_intersect(psx_scene *handle, const psx_ray_query *queries, uint32_t count, psx_ray_hit *hits) {
    Scene *s = reinterpret_cast<Scene *>(handle);
    if (!s || (count && (!queries || !hits))) return PSX_ERROR_INVALID_ARGUMENT;
    if (count == 0u) return PSX_OK;
    return guarded([&] {
        hip_check(hipSetDevice(s->backend->device));
        DeviceBuffer q, h;
        q.allocate(static_cast<size_t>(count) * sizeof(psx_ray_query));
        h.allocate(static_cast<size_t>(count) * sizeof(psx_ray_hit));
        hip_check(hipMemcpy(q.as<void>(), queries, static_cast<size_t>(count) * sizeof(psx_ray_query),
                            hipMemcpyHostToDevice));
        const int rc = psx_scene_intersect_device(handle, q.as<psx_ray_query>(), count, h.as<psx_ray_hit>());
        if (rc != PSX_OK) return rc;
        hip_check(hipMemcpy(hits, h.as<void>(), static_cast<size_t>(count) * sizeof(psx_ray_hit),
                            hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

int psx_scene_intersect_device_noalloc(psx_scene *handle, const psx_ray_query *queries, uint32_t count,
                                       psx_ray_hit *hits) {
    Scene *s = reinterpret_cast<Scene *>(handle);
    if (!s || (count && (!queries || !hits)) || count > PSX_INTERSECT_NOALLOC_MAX_QUERIES)
        return PSX_ERROR_INVALID_ARGUMENT;
    if (count == 0u) return PSX_OK;
    return guarded([&] {
        hip_check(hipSetDevice(s->backend->device));
        hip_check(launch_intersect_noalloc(s->view, queries, count, hits, s->query_stream));
        hip_check(hipStreamSynchronize(s->query_stream));
        return static_cast<int>(PSX_OK);
    });
}

uint32_t psx_probe_input_width(uint32_t function) {
    static const uint32_t w[] = {6u, 1u, 21u, 47u, 14u, 15u, 13u, 59u, 21u, 32u};
    return function < sizeof(w) / sizeof(w[0]) ? w[function] : 0u;
}

uint32_t psx_probe_output_width(uint32_t function) {
    static const uint32_t w[] = {3u, 1u, 4u, 4u, 10u, 3u, 4u, 4u, 10u, 24u};
    return function < sizeof(w) / sizeof(w[0]) ? w[function] : 0u;
}

int psx_backend_probe(psx_backend *handle, uint32_t function, const float *inputs, uint32_t count, float *outputs) {
    Backend *b = reinterpret_cast<Backend *>(handle);
    const uint32_t wi = psx_probe_input_width(function), wo = psx_probe_output_width(function);
    if (!b || wi == 0u || (count && (!inputs || !outputs))) return PSX_ERROR_INVALID_ARGUMENT;
    if (count == 0u) return PSX_OK;
    return guarded([&] {
        hip_check(hipSetDevice(b->device));
        DeviceBuffer in, out;
        in.allocate(static_cast<size_t>(count) * wi * sizeof(float));
        out.allocate(static_cast<size_t>(count) * wo * sizeof(float));
        hip_check(hipMemcpy(in.as<void>(), inputs, static_cast<size_t>(count) * wi * sizeof(float),
                            hipMemcpyHostToDevice));
        hip_check(launch_probe(function, in.as<float>(), wi, count, out.as<float>(), wo, nullptr));
        hip_check(hipDeviceSynchronize());
        hip_check(hipMemcpy(outputs, out.as<void>(), static_cast<size_t>(count) * wo * sizeof(float),
                            hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

}  // extern "C"

// ---- render server ---------------------------------------------------------
int psx_server_create(psx_backend *b, const psx_server_options *options, psx_server **out) {
    if (!b || !out) return PSX_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    psx_server_options o = options ? *options : psx_server_options{};
    if (o.path_capacity == 0u) o.path_capacity = 8388608u;
    o.path_capacity = (o.path_capacity + 63u) & ~63u;
    if (o.streams == 0u) o.streams = 6u;
    if (o.streams < 2u || o.streams > 8u || o.path_capacity > (1u << 30)) return PSX_ERROR_INVALID_ARGUMENT;
    if (o.hbm_budget == 0u) o.hbm_budget = 200ull << 30;
    if (o.viewport_max_pixels == 0u) o.viewport_max_pixels = 256000u;
    if (o.viewport_max_samples == 0u) o.viewport_max_samples = 4u;
    if (o.target_batch_ms == 0u) o.target_batch_ms = 12u;
    return guarded([&] {
        hip_check(hipSetDevice(b->device));
        auto server = std::make_unique<psx_server>();
        server->backend = b;
        server->options = o;
        hip_check(hipStreamCreateWithFlags(&server->control, hipStreamNonBlocking));
        const size_t lanes = wavefront_lanes();
        for (uint32_t i = 0; i < o.streams; ++i) {
            auto a = std::make_unique<Arena>();
            a->state.allocate(o.path_capacity, 0u, kMaxIterations, false);
            a->item_capacity = 256u;
            a->items.allocate(a->item_capacity * sizeof(BatchItem));
            a->spill.allocate(lanes * kTraversalStackDepth * sizeof(uint32_t));
            a->counters.allocate(sizeof(Counters));
            hip_check(hipHostMalloc(reinterpret_cast<void **>(&a->host_items), a->item_capacity * sizeof(BatchItem),
                                    hipHostMallocDefault));
            hip_check(hipHostMalloc(reinterpret_cast<void **>(&a->host_counters), sizeof(Counters),
                                    hipHostMallocDefault));
            std::memset(a->host_counters, 0, sizeof(Counters));
            hip_check(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
            hip_check(hipEventCreate(&a->start));
            hip_check(hipEventCreate(&a->done));
            for (hipEvent_t &e : a->ring) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            hip_check(hipHostMalloc(reinterpret_cast<void **>(&a->host_queues),
                                    (kMaxIterations + 1u) * sizeof(WavefrontCounters), hipHostMallocDefault));
            std::memset(a->host_queues, 0, (kMaxIterations + 1u) * sizeof(WavefrontCounters));
            // the upper half of the streams takes viewport batches only
            a->viewport_only = i >= o.streams - std::max(o.streams / 2u, 1u);
            a->bytes = static_cast<size_t>(o.path_capacity) * 288u + a->items.bytes() + a->spill.bytes();
            server->fixed_bytes += a->bytes;
            server->arenas.push_back(std::move(a));
        }
        if (server->fixed_bytes > o.hbm_budget) return static_cast<int>(PSX_ERROR_OUT_OF_MEMORY);
        server->stats.hbm_in_use = server->stats.peak_hbm = server->fixed_bytes;
        psx_server *raw = server.get();
        raw->worker = std::thread([raw] { raw->run(); });
        for (uint32_t i = 0; i < kEditThreads; ++i) raw->editors.emplace_back([raw] { raw->edit_loop(); });
        *out = server.release();
        return static_cast<int>(PSX_OK);
    });
}

void psx_server_destroy(psx_server *server) {
    ThreadLocalCaptureMode mode;
    delete server;
}

int psx_server_submit(psx_server *server, psx_scene *handle, const psx_render_settings *rs, uint32_t first,
                      uint32_t count, uint32_t total, uint64_t *out_id) {
    Scene *scene = reinterpret_cast<Scene *>(handle);
    if (!server || !scene || !out_id || !valid_settings(rs)) return PSX_ERROR_INVALID_ARGUMENT;
    if (scene->backend != server->backend) return PSX_ERROR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < rs->pass_count; ++i)
        if (light_pass_slot(rs->passes[i]) >= 0) return PSX_ERROR_UNSUPPORTED;
    if (count == 0u || total == 0u || static_cast<uint64_t>(first) + count > total) return PSX_ERROR_RANGE;
    const uint32_t ww = rs->window_width ? rs->window_width : rs->width;
    const uint32_t wh = rs->window_height ? rs->window_height : rs->height;
    const uint64_t padded = static_cast<uint64_t>((ww + 7u) / 8u) * ((wh + 7u) / 8u) * 64u;
    if (padded > server->options.path_capacity) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        auto r = std::make_shared<ServerRequest>();
        r->scene = scene;
        keep_settings(r->film, *rs);
        r->first = first;
        r->count = count;
        r->total = total;
        {
            std::lock_guard<std::mutex> lock(server->mutex);
            if (server->worker_error != PSX_OK) return server->worker_error;
            r->id = server->next_id++;
            r->submit_ms = server->now_ms();
            server->requests.emplace(r->id, r);
            ServerOp op;
            op.request = r;
            server->ops.push_back(std::move(op));
            server->stats.submitted += 1u;
            *out_id = r->id;
        }
        server->wake.notify_all();
        return static_cast<int>(PSX_OK);
    });
}

int psx_server_update_parameters(psx_server *server, psx_scene *handle, uint32_t material, const float *p,
                                 uint32_t count) {
    Scene *scene = reinterpret_cast<Scene *>(handle);
    if (!server || !scene || !p || material >= scene->materials.size()) return PSX_ERROR_INVALID_ARGUMENT;
    if (count != scene->materials[material].params.size()) return PSX_ERROR_INVALID_ARGUMENT;
    return guarded([&] {
        {
            std::lock_guard<std::mutex> lock(server->mutex);
            if (server->worker_error != PSX_OK) return server->worker_error;
            ServerOp op;
            op.scene = scene;
            op.material = material;
            op.params.assign(p, p + count);
            server->ops.push_back(std::move(op));
        }
        server->wake.notify_all();
        return static_cast<int>(PSX_OK);
    });
}

namespace {
std::shared_ptr<ServerRequest> find_request(psx_server *server, uint64_t id) {
    auto it = server->requests.find(id);
    return it == server->requests.end() ? nullptr : it->second;
}
}  // namespace

int psx_server_status(psx_server *server, uint64_t id, psx_request_status *out) {
    if (!server || !out) return PSX_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(server->mutex);
    auto r = find_request(server, id);
    if (!r) return PSX_ERROR_INVALID_ARGUMENT;
    out->state = r->state;
    out->samples_done = r->done;
    out->batches = r->batches;
    out->viewport = r->viewport ? 1u : 0u;
    out->submit_ms = r->submit_ms;
    out->admit_ms = r->admit_ms;
    out->first_launch_ms = r->first_launch_ms;
    out->done_ms = r->done_ms;
    return PSX_OK;
}

int psx_server_wait(psx_server *server, uint64_t id, double timeout_ms) {
    if (!server) return PSX_ERROR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(server->mutex);
    auto r = find_request(server, id);
    if (!r) return PSX_ERROR_INVALID_ARGUMENT;
    auto ended = [&] { return r->state >= PSX_REQUEST_DONE || server->worker_error != PSX_OK; };
    if (timeout_ms < 0.0) {
        server->finished.wait(lock, ended);
    } else if (!server->finished.wait_for(lock, std::chrono::duration<double, std::milli>(timeout_ms), ended)) {
        return PSX_ERROR_INCOMPLETE;
    }
    if (r->state == PSX_REQUEST_DONE) return PSX_OK;
    if (r->state == PSX_REQUEST_CANCELLED) return PSX_ERROR_CANCELLED;
    if (r->error != PSX_OK) {
        g_last_error = r->message;
        return r->error;
    }
    g_last_error = server->worker_message;
    return server->worker_error != PSX_OK ? server->worker_error : static_cast<int>(PSX_ERROR_DEVICE);
}

int psx_server_read_film(psx_server *server, uint64_t id, float *combined, float *normal, float *albedo,
                         uint32_t *counts) {
    if (!server) return PSX_ERROR_INVALID_ARGUMENT;
    std::shared_ptr<ServerRequest> r;
    {
        std::lock_guard<std::mutex> lock(server->mutex);
        r = find_request(server, id);
        if (!r) return PSX_ERROR_INVALID_ARGUMENT;
        if (r->state != PSX_REQUEST_DONE) return PSX_ERROR_INCOMPLETE;
    }
    return guarded([&] {
        hip_check(hipSetDevice(server->backend->device));
        const Session &f = r->film;
        const size_t n = static_cast<size_t>(f.window_w) * f.window_h;
        if (combined) hip_check(hipMemcpy(combined, f.combined.as<void>(), n * 16u, hipMemcpyDeviceToHost));
        if (normal) hip_check(hipMemcpy(normal, f.normal.as<void>(), n * 16u, hipMemcpyDeviceToHost));
        if (albedo) hip_check(hipMemcpy(albedo, f.albedo.as<void>(), n * 16u, hipMemcpyDeviceToHost));
        if (counts) hip_check(hipMemcpy(counts, f.sample_count.as<void>(), n * 4u, hipMemcpyDeviceToHost));
        return static_cast<int>(PSX_OK);
    });
}

int psx_server_write_passes(psx_server *server, uint64_t id, const psx_sink *sink) {
    if (!server) return PSX_ERROR_INVALID_ARGUMENT;
    std::shared_ptr<ServerRequest> r;
    {
        std::lock_guard<std::mutex> lock(server->mutex);
        r = find_request(server, id);
        if (!r) return PSX_ERROR_INVALID_ARGUMENT;
        if (r->state != PSX_REQUEST_DONE) return PSX_ERROR_INCOMPLETE;
    }
    return guarded([&] {
        hip_check(hipSetDevice(server->backend->device));
        return write_passes(r->film, sink);
    });
}

int psx_server_cancel(psx_server *server, uint64_t id) {
    if (!server) return PSX_ERROR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lock(server->mutex);
        auto r = find_request(server, id);
        if (!r) return PSX_ERROR_INVALID_ARGUMENT;
        r->cancel.store(true);
    }
    server->wake.notify_all();
    return PSX_OK;
}

int psx_server_release(psx_server *server, uint64_t id) {
    if (!server) return PSX_ERROR_INVALID_ARGUMENT;
    std::shared_ptr<ServerRequest> r;
    {
        std::lock_guard<std::mutex> lock(server->mutex);
        auto it = server->requests.find(id);
        if (it == server->requests.end()) return PSX_ERROR_INVALID_ARGUMENT;
        if (it->second->state < PSX_REQUEST_DONE || it->second->in_flight) return PSX_ERROR_INCOMPLETE;
        r = std::move(it->second);
        server->requests.erase(it);
    }
    // the film is freed here unless the worker still holds the request
    return guarded([&] {
        hip_check(hipSetDevice(server->backend->device));
        r.reset();
        return static_cast<int>(PSX_OK);
    });
}

int psx_server_stats(psx_server *server, psx_serving_stats *out) {
    if (!server || !out) return PSX_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(server->mutex);
    *out = server->stats;
    out->hbm_in_use = server->fixed_bytes + server->film_bytes + server->resident_bytes();
    return PSX_OK;
}
