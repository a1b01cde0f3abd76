// contract_driver.cpp - drives the render path through include/psx_contract.hpp,
// the C++ mirror of the reference's RendererBackend / RenderSession /
// OutputSink contract (include/psycles/contract/render.h:136-181), as a host
// written against that contract would.
//
// Without a device it checks that the backend reports the failure (the
// constructor throws) and exits 0 after printing "no-device".  With a device
// it compiles a lit floor, renders two progressive sample ranges into a
// collecting sink, and checks the contract's error behaviour: a null scene
// plus diagnostics for an invalid snapshot, a null session for invalid
// settings, false from render_samples for a bad range and for a changed
// total.  Prints "ok" on success; any failure prints its reason and exits 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "../include/psx_contract.hpp"

namespace {

int failures = 0;

void expect(bool condition, const char *what) {
    if (!condition) {
        std::printf("FAILED: %s\n", what);
        ++failures;
    }
}

class CollectingSink final : public psx::OutputSink {
  public:
    int begins = 0, ends = 0;
    bool cancelled = false;
    std::map<uint32_t, std::vector<float>> passes;
    std::map<uint32_t, uint32_t> channels;

    void begin(const psx_render_settings &) override { ++begins; }
    void write(const psx::PassTile &tile) override {
        passes[tile.pass].assign(tile.pixels.begin(), tile.pixels.end());
        channels[tile.pass] = tile.channels;
    }
    void end(bool was_cancelled) override {
        ++ends;
        cancelled = was_cancelled;
    }
};

void identity(float *m) {
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

}  // namespace

int main() {
    if (psx_device_count() <= 0) {
        bool threw = false;
        try {
            psx::RendererBackend backend{0};
        } catch (const std::runtime_error &) {
            threw = true;
        }
        expect(threw, "backend creation without a device must fail");
        std::printf(failures ? "failed\n" : "no-device\n");
        return failures ? 1 : 0;
    }

    // a floor quad (grey diffuse) under an emissive quad facing down
    const float floor_p[] = {-2, 0, -2, 2, 0, -2, 2, 0, 2, -2, 0, 2};
    const uint32_t quad_t[] = {0, 2, 1, 0, 3, 2};
    const float light_p[] = {-0.5f, 1.5f, -0.5f, 0.5f, 1.5f, -0.5f, 0.5f, 1.5f, 0.5f, -0.5f, 1.5f, 0.5f};
    const uint32_t light_t[] = {0, 1, 2, 0, 2, 3};
    const psx_closure_instr diffuse{PSX_CLOSURE_DIFFUSE, 0, 0, 0, 0, 0, 3, 0};
    const psx_closure_instr emission{PSX_CLOSURE_EMISSION, 0, 0, 0, 0, 3, 0, 0};
    const float grey[] = {0.7f, 0.7f, 0.7f, 0.0f};
    const float white[] = {1.0f, 1.0f, 1.0f, 8.0f};
    psx_material materials[2] = {{&diffuse, 1, 0, grey, 4, PSX_EMISSION_SAMPLING_AUTOMATIC},
                                 {&emission, 1, 0, white, 4, PSX_EMISSION_SAMPLING_AUTOMATIC}};
    const uint32_t slot0 = 0, slot1 = 1;
    psx_mesh meshes[2] = {};
    meshes[0] = {floor_p, 4, quad_t, 2, nullptr, nullptr, &slot0, 1, nullptr};
    meshes[1] = {light_p, 4, light_t, 2, nullptr, nullptr, &slot1, 1, nullptr};
    psx_instance instances[2] = {};
    for (uint32_t i = 0; i < 2; ++i) {
        instances[i].mesh = i;
        instances[i].visibility_mask = PSX_VISIBILITY_ALL;
        identity(instances[i].transform);
    }
    psx_scene_desc desc{};
    desc.abi_version = PSX_ABI_VERSION;
    desc.materials = materials;
    desc.material_count = 2;
    desc.meshes = meshes;
    desc.mesh_count = 2;
    desc.instances = instances;
    desc.instance_count = 2;
    identity(desc.camera.transform);
    desc.camera.transform[13] = 0.8f;   // camera at (0, 0.8, 4) looking down -z
    desc.camera.transform[14] = 4.0f;
    desc.camera.field_of_view = 0.8f;
    desc.camera.horizontal_field_of_view = 0.8f;
    desc.camera.sensor_fit = 1;
    desc.camera.near_clip = 1e-3f;
    desc.camera.far_clip = 100.0f;
    desc.world_visibility_mask = PSX_VISIBILITY_ALL;

    psx::RendererBackend backend{0};
    psx::SceneCompilation compilation = backend.compile_scene(desc);
    expect(compilation.ok(), "the floor scene compiles");
    if (!compilation.ok()) return 1;

    psx_render_settings settings;
    psx_render_settings_default(&settings);
    settings.width = settings.window_width = 48;
    settings.height = settings.window_height = 32;
    const uint32_t passes[] = {PSX_PASS_COMBINED, PSX_PASS_NORMAL, PSX_PASS_ALBEDO, PSX_PASS_SAMPLE_COUNT};
    const uint32_t chans[] = {4, 3, 3, 1};
    settings.passes = passes;
    settings.pass_channels = chans;
    settings.pass_count = 4;
    auto session = backend.create_session(*compilation.scene, settings);
    expect(session != nullptr, "a session for valid settings");
    if (!session) return 1;

    CollectingSink sink;
    expect(session->render_samples({0, 2, 0, 4}, sink), "samples 0-1 of 4");
    expect(session->render_samples({2, 2, 0, 4}, sink), "samples 2-3 of 4");
    expect(sink.begins == 2 && sink.ends == 2 && !sink.cancelled, "begin / end once per call");
    expect(sink.passes.size() == 4, "every requested pass written");
    const auto &count = sink.passes[PSX_PASS_SAMPLE_COUNT];
    const auto &combined = sink.passes[PSX_PASS_COMBINED];
    expect(count.size() == 48u * 32u && combined.size() == 48u * 32u * 4u, "pass shapes");
    bool counts_ok = !count.empty(), lit = false;
    for (float c : count) counts_ok = counts_ok && c == 4.0f;
    for (size_t i = 0; i < combined.size(); ++i) {
        if (!std::isfinite(combined[i])) counts_ok = false;
        if (i % 4 != 3 && combined[i] > 0.0f) lit = true;
    }
    expect(counts_ok, "four samples per pixel, finite values");
    expect(lit, "the floor is lit by the emitter");

    // the contract's error behaviour
    CollectingSink other;
    expect(!session->render_samples({0, 8, 0, 4}, other), "a range past total is rejected");
    expect(!session->render_samples({4, 1, 0, 5}, other), "a changed total is rejected");
    psx_render_settings bad = settings;
    bad.width = 0;
    expect(backend.create_session(*compilation.scene, bad) == nullptr, "no session for an empty extent");
    psx_scene_desc broken = desc;
    psx_instance orphan = instances[0];
    orphan.mesh = 7;
    broken.instances = &orphan;
    broken.instance_count = 1;
    psx::SceneCompilation failed = backend.compile_scene(broken);
    expect(!failed.ok() && !failed.diagnostics.empty(), "a missing mesh gives a null scene and a diagnostic");

    const float brighter[] = {1.0f, 1.0f, 1.0f, 16.0f};
    expect(compilation.scene->update_parameters(1, std::span<const float>{brighter, 4}), "parameter rebind");
    expect(!compilation.scene->update_parameters(1, std::span<const float>{brighter, 3}), "structure change refused");

    std::printf(failures ? "failed\n" : "ok\n");
    return failures ? 1 : 0;
}
