// psx_rtc.h - surface programs generated per closure-leaf layout and compiled
// at run time with hipRTC.
//
// The reference compiles each material's SurfaceProgram into device code
// once per structure signature and keeps socket values in parameter buffers
// (docs/architecture.md:63-110; include/psycles/compiler/surface_program.h:
// 512-568), so a parameter-only edit rebinds data without recompiling.  Here
// a material's closure tree is flattened to leaves (operation, mix depth,
// mix-factor, colour and strength slots of its parameter block, mix sides);
// materials with the same leaf layout share one generated program, and the
// parameter blocks stay in HBM.  A program is straight-line HIP that performs
// the flat-leaf interpreter's operations for that layout in the same order,
// so films are bit-identical to the interpreter's.  One generated translation
// unit holds a scene's programs, a switch over them and one extern "C" shade
// kernel that runs the wavefront shade step with them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "psx_device.h"

namespace psx {

using LeafLayout = std::vector<Leaf>;

// Index of `layout` in `layouts`, appended when new: the material's program.
uint32_t program_index(std::vector<LeafLayout> &layouts, const LeafLayout &layout);

// Shade-kernel variants: bit 0 light-pass sums, bit 1 lights in the scene.
constexpr uint32_t kProgramLightPasses = 1u, kProgramLights = 2u;
constexpr const char *kGeneratedShadeName = "psx_generated_shade";

// The generated translation unit for `layouts` (program i = layouts[i]).
std::string surface_program_source(const std::vector<LeafLayout> &layouts, uint32_t variant);

struct RtcResult {
    bool ok = false;
    std::string log;
    std::vector<char> code;  // code object for hipModuleLoadData
};

// hipRTC compilation of a generated source for `arch` ("gfx950", or a
// device's gcnArchName) with the flags of the compiled-in kernels; needs no
// device.
RtcResult compile_surface_program(const std::string &source, const std::string &arch);

struct GeneratedShade {
    hipFunction_t fn = nullptr;
    uint32_t blocks = 0;  // resident blocks of kShadeBlock threads on the device
};

// Generated shade kernels loaded on one device, keyed by their source; a
// scene whose layouts and variant match an earlier one reuses its module.
class ProgramCache {
  public:
    ProgramCache() = default;
    ProgramCache(const ProgramCache &) = delete;
    ProgramCache &operator=(const ProgramCache &) = delete;
    ~ProgramCache();
    // Compiles and loads on first use.  Throws std::runtime_error carrying
    // the compiler log when the program does not compile or load.
    GeneratedShade get(const std::vector<LeafLayout> &layouts, uint32_t variant, int device);
    uint32_t compiled() const;

  private:
    struct Entry {
        std::string source;
        hipModule_t module = nullptr;
        GeneratedShade shade;
    };
    mutable std::mutex mutex_;
    std::vector<Entry> entries_;
};

}  // namespace psx
