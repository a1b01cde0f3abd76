// psx_rtc.cpp - generation and hipRTC compilation of surface programs (see
// psx_rtc.h).
#include "psx_rtc.h"

#include <hip/hiprtc.h>

#include <cstring>
#include <sstream>
#include <stdexcept>

namespace psx {

// Device headers the generated source includes, embedded into the library
// at build time (psycles_amd/build.py writes them from psycles_amd/csrc).
extern const unsigned kEmbeddedCount;
extern const char *const kEmbeddedNames[];
extern const char *const kEmbeddedSources[];

namespace {

constexpr uint32_t kShadeThreads = 256u;  // kShadeBlock of the wavefront kernels

bool same_leaf(const Leaf &a, const Leaf &b) { return std::memcmp(&a, &b, sizeof(Leaf)) == 0; }

// Appends a closure the way the interpreter does: slot writes through
// constant indices, which keep the set in registers.
void emit_append(std::ostringstream &o, const char *kind, const char *weight, const char *sw,
                 const std::string &roughness, const char *color) {
    o << "#pragma unroll\n"
      << "            for (uint32_t c = 0; c < kMaxClosures; ++c) {\n"
      << "                if (c == cs.count) {\n"
      << "                    cs.kind[c] = " << kind << ";\n"
      << "                    cs.weight[c] = " << weight << ";\n"
      << "                    cs.sw[c] = " << sw << ";\n"
      << "                    cs.roughness[c] = " << roughness << ";\n"
      << "                    cs.color[c] = " << color << ";\n"
      << "                }\n"
      << "            }\n"
      << "            cs.count += 1u;\n";
}

// One leaf's statements: the body of the interpreter's leaf loop
// (material_closures in psx_path.h) with the leaf's fields as constants.
void emit_leaf(std::ostringstream &o, const Leaf &leaf, uint32_t index) {
    const bool emission = leaf.op == kLeafEmission, diffuse = leaf.op == kLeafDiffuse,
               transparent = leaf.op == kLeafTransparent;
    o << "    {  // leaf " << index << ": "
      << (emission ? "emission" : diffuse ? "diffuse" : transparent ? "transparent" : "inert") << "\n";
    if (!emission && !diffuse && !transparent) {
        o << "    }\n";
        return;
    }
    o << "        float w = 1.0f;\n";
    for (uint32_t d = 0; d < leaf.depth && d < kMaxLeafDepth; ++d) {
        const bool second = ((leaf.second_mask >> d) & 1u) != 0u;
        o << "        {\n            const float f = clampf(p[" << leaf.factor[d] << "u], 0.0f, 1.0f);\n"
          << "            w = w * " << (second ? "f" : "(1.0f - f)") << ";\n        }\n";
    }
    o << "        const f3 color = V(p[" << leaf.color << "u], p[" << leaf.color + 1u << "u], p[" << leaf.color + 2u
      << "u]);\n";
    if (emission) {
        o << "        cs.emission = cs.emission + (color * p[" << leaf.aux << "u]) * w;\n";
    } else if (diffuse) {
        o << "        f3 wt = vmax(color * w, V(0.0f, 0.0f, 0.0f));\n"
          << "        if (!(avg3(wt) >= 1.0e-5f)) wt = V(0.0f, 0.0f, 0.0f);\n"
          << "        const float sw = fabsf(avg3(vmax(wt, V(0.0f, 0.0f, 0.0f))));\n"
          << "        wt = vmax(wt, V(0.0f, 0.0f, 0.0f));\n"
          << "        if (sw >= 1.0e-5f && cs.count < kMaxClosures) {\n";
        emit_append(o, "KIND_DIFFUSE", "wt", "sw", "p[" + std::to_string(leaf.aux) + "u]", "color");
        o << "        }\n";
    } else {
        o << "        const f3 wt = color * w;\n"
          << "        const float sw = fabsf(avg3(wt));\n"
          << "        if (sw >= 1.0e-5f && cs.count < kMaxClosures) {\n";
        emit_append(o, "KIND_TRANSPARENT", "wt", "sw", "0.0f", "V(1.0f, 1.0f, 1.0f)");
        o << "        }\n";
    }
    o << "    }\n";
}

}  // namespace

uint32_t program_index(std::vector<LeafLayout> &layouts, const LeafLayout &layout) {
    for (size_t i = 0; i < layouts.size(); ++i) {
        const LeafLayout &l = layouts[i];
        if (l.size() != layout.size()) continue;
        bool same = true;
        for (size_t j = 0; j < l.size() && same; ++j) same = same_leaf(l[j], layout[j]);
        if (same) return static_cast<uint32_t>(i);
    }
    layouts.push_back(layout);
    return static_cast<uint32_t>(layouts.size() - 1u);
}

std::string surface_program_source(const std::vector<LeafLayout> &layouts, uint32_t variant) {
    std::ostringstream o;
    o << "// generated surface programs: " << layouts.size() << " closure-leaf layout(s), shade variant "
      << variant << "\n"
      << "#define PSX_GENERATED_SURFACE_PROGRAMS 1\n"
      << "#include \"psx_wavefront_kernels.h\"\n\n"
      << "namespace psx {\nnamespace {\n\n";
    for (size_t i = 0; i < layouts.size(); ++i) {
        o << "__device__ __forceinline__ void surface_program_" << i << "(const float *p, ClosureSet &cs) {\n";
        if (layouts[i].empty()) o << "    (void)p;\n    (void)cs;\n";
        for (size_t j = 0; j < layouts[i].size(); ++j) emit_leaf(o, layouts[i][j], static_cast<uint32_t>(j));
        o << "}\n\n";
    }
    o << "__device__ void generated_closures(uint32_t program, const float *p, ClosureSet &cs) {\n"
      << "    switch (program) {\n";
    for (size_t i = 0; i < layouts.size(); ++i)
        o << "    case " << i << "u: surface_program_" << i << "(p, cs); return;\n";
    o << "    default: return;\n    }\n}\n\n}  // namespace\n}  // namespace psx\n\n";
    const char *lp = (variant & kProgramLightPasses) ? "true" : "false";
    const char *lights = (variant & kProgramLights) ? "true" : "false";
    o << "extern \"C\" __global__ __launch_bounds__(" << kShadeThreads << ", 2) void " << kGeneratedShadeName
      << "(psx::SceneView s, psx::KernelParams k, const float4 *sobol, psx::WavefrontView w, uint32_t it) {\n"
      << "    psx::shade_launch<" << lp << ", " << lights << ">(s, k, sobol, w, it);\n}\n";
    return o.str();
}

RtcResult compile_surface_program(const std::string &source, const std::string &arch) {
    RtcResult r;
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, source.c_str(), "psx_surface_programs.hip", static_cast<int>(kEmbeddedCount),
                            kEmbeddedSources, kEmbeddedNames) != HIPRTC_SUCCESS) {
        r.log = "hiprtcCreateProgram failed";
        return r;
    }
    // the flags of the compiled-in kernels (psycles_amd/build.py): the
    // floating-point contract and packed FP32 disabled
    const std::string arch_flag = "--offload-arch=" + arch;
    const char *options[] = {arch_flag.c_str(),
                             "-O3",
                             "-std=c++17",
                             "-ffp-contract=off",
                             "-fno-fast-math",
                             "-fhip-fp32-correctly-rounded-divide-sqrt",
                             "-Xclang",
                             "-target-feature",
                             "-Xclang",
                             "-packed-fp32-ops"};
    const hiprtcResult compiled =
        hiprtcCompileProgram(prog, static_cast<int>(sizeof(options) / sizeof(options[0])), options);
    size_t log_size = 0;
    if (hiprtcGetProgramLogSize(prog, &log_size) == HIPRTC_SUCCESS && log_size > 1u) {
        r.log.resize(log_size);
        if (hiprtcGetProgramLog(prog, &r.log[0]) != HIPRTC_SUCCESS) r.log.clear();
        while (!r.log.empty() && r.log.back() == '\0') r.log.pop_back();
    }
    if (compiled == HIPRTC_SUCCESS) {
        size_t code_size = 0;
        if (hiprtcGetCodeSize(prog, &code_size) == HIPRTC_SUCCESS && code_size > 0u) {
            r.code.resize(code_size);
            r.ok = hiprtcGetCode(prog, r.code.data()) == HIPRTC_SUCCESS;
        }
    } else if (r.log.empty()) {
        r.log = hiprtcGetErrorString(compiled);
    }
    (void)hiprtcDestroyProgram(&prog);
    return r;
}

ProgramCache::~ProgramCache() {
    for (Entry &e : entries_)
        if (e.module) (void)hipModuleUnload(e.module);
}

uint32_t ProgramCache::compiled() const {
    std::lock_guard<std::mutex> lock(mutex_);
    return static_cast<uint32_t>(entries_.size());
}

GeneratedShade ProgramCache::get(const std::vector<LeafLayout> &layouts, uint32_t variant, int device) {
    const std::string source = surface_program_source(layouts, variant);
    std::lock_guard<std::mutex> lock(mutex_);
    for (const Entry &e : entries_)
        if (e.source == source) return e.shade;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) throw std::runtime_error("hipGetDeviceProperties failed");
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    const RtcResult r = compile_surface_program(source, arch);
    if (!r.ok) throw std::runtime_error("surface program compilation failed: " + r.log);
    Entry e;
    e.source = source;
    if (hipModuleLoadData(&e.module, r.code.data()) != hipSuccess)
        throw std::runtime_error("hipModuleLoadData failed for a generated surface program");
    if (hipModuleGetFunction(&e.shade.fn, e.module, kGeneratedShadeName) != hipSuccess) {
        (void)hipModuleUnload(e.module);
        throw std::runtime_error("generated shade kernel not found in its module");
    }
    int per_cu = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, e.shade.fn, static_cast<int>(kShadeThreads), 0) !=
            hipSuccess ||
        per_cu <= 0) {
        (void)hipModuleUnload(e.module);
        throw std::runtime_error("occupancy query failed for a generated shade kernel");
    }
    e.shade.blocks = static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(prop.multiProcessorCount);
    entries_.push_back(std::move(e));
    return entries_.back().shade;
}

}  // namespace psx
