"""Blender bundle loader (psycles_amd.blender) on the CPU.

The settings round trip restates tests/test_blender_import.cpp:62-860
(test_integrator_settings_round_trip): the same scene.json values and the
same expected render size, seed, filter, integrator, light and world
fields.  The geometry tests write a SceneSnapshot as a v1 or v2 bundle,
load it back and render both with the oracle: bit-identical films pin the
whole mapping (matrices, slots, lights, camera, node-tree lowering).
"""
import copy
import json
import math
import os

import numpy as np
import pytest

import blender_bundle as bb
import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.blender import load_blender_scene_bundle, visibility_mask


def emission_tree(color=(1.0, 0.5, 0.25), strength=3.0):
    return {"name": "Emissive", "nodes": [bb.node("Emission", "EMISSION",
                                                  [bb.socket("Color", list(color) + [1.0], "NodeSocketColor"),
                                                   bb.socket("Strength", strength)])],
            "links": [], "surface_root": {"node": "Emission", "socket": "Emission"},
            "volume_root": None, "displacement_root": None}


def reference_scene():
    """scene.json of test_integrator_settings_round_trip, with the volume
    material's tree replaced by a surface emission (volumes are outside
    this subset; `test_volume_material_is_reported` keeps the original)."""
    identity = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]
    return {
        "schema": "psycles.blender-scene.v1", "images": [], "node_groups": [],
        "materials": [{"name": "Raw Emissive", "use_bump_map_correction": False, "emission_sampling": "BACK",
                       "volume_sampling": "EQUIANGULAR", "cycles_sync": {"shader_index": 6},
                       "node_tree": emission_tree()}],
        "geometries": [], "instances": [],
        "lights": [{"name": "Key", "type": "AREA", "transform": identity, "color": [1, 1, 1], "energy": 10,
                    "shape": "DISK", "size": 1.3, "size_y": 0.25, "use_multiple_importance_sampling": False,
                    "max_bounces": 13, "cast_shadow": False,
                    "visibility": {"camera": False, "diffuse": True, "glossy": False, "transmission": True,
                                   "shadow": True, "volume_scatter": False},
                    "is_shadow_catcher": True,
                    "cycles_sync": {"shader_index": 5, "object_index": 9, "light_group": 2},
                    "node_tree": None}],
        "world": {"name": "World", "color": [0.05, 0.05, 0.05], "sampling_method": "MANUAL",
                  "sample_map_resolution": 2048, "max_bounces": 23, "use_shadows": False,
                  "visibility": {"camera": True, "diffuse": False, "glossy": True, "transmission": False,
                                 "shadow": True, "volume_scatter": False},
                  "cycles_sync": {"shader_index": 3, "object_index": 12, "light_group": 4}},
        "world_environment": None,
        "render": {"width": 480, "height": 270, "percentage": 50, "transparent": True,
                   "pixel_filter_type": "GAUSSIAN", "filter_width": 2.25, "pass_alpha_threshold": 0.375,
                   "cycles": {"samples": 37, "seed": 305419896, "effective_seed": 2271560481,
                              "use_adaptive_sampling": False, "use_denoising": False, "max_bounces": 11,
                              "min_light_bounces": 2, "diffuse_bounces": 3, "glossy_bounces": 4,
                              "transmission_bounces": 5, "volume_bounces": 6, "min_transparent_bounces": 1,
                              "transparent_max_bounces": 9, "sample_clamp_direct": 1.25,
                              "sample_clamp_indirect": 2.5, "blur_glossy": 0.75, "film_exposure": 0.75,
                              "light_sampling_threshold": 0.125, "caustics_reflective": False,
                              "caustics_refractive": True, "use_light_tree": False,
                              "direct_light_sampling_type": "NEXT_EVENT_ESTIMATION"}},
        "camera": {"name": "Camera", "type": "PERSP", "transform": identity, "angle": 0.75, "angle_x": 0.9,
                   "angle_y": 0.6, "sensor_fit": "AUTO", "ortho_scale": 1.0, "shift_x": 0.0, "shift_y": 0.0,
                   "clip_start": 0.01, "clip_end": 100.0, "lens": 50.0,
                   "dof": {"enabled": False, "focus_distance": 10.0, "fstop": 2.8, "blades": 0,
                           "rotation": 0.0, "ratio": 1.0}}}


def write(directory, scene, geometry=b"PSYGEO1\0"):
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "scene.json"), "w") as f:
        json.dump(scene, f)
    with open(os.path.join(directory, "geometry.bin"), "wb") as f:
        f.write(geometry)
    return directory


def errors(imported):
    return [d.message for d in imported.diagnostics if d.severity == "error"]


def test_integrator_settings_round_trip(tmp_path):
    imported = load_blender_scene_bundle(write(tmp_path / "b", reference_scene()))
    assert imported.ok(), errors(imported)
    assert (imported.width, imported.height, imported.samples) == (240, 135, 37)
    assert imported.seed == 2271560481
    s = imported.settings
    assert s.transparent_background == 1
    assert s.pixel_filter == abi.FILTER_GAUSSIAN and s.filter_width == pytest.approx(2.25, abs=1e-6)
    assert s.pass_alpha_threshold == pytest.approx(0.375, abs=1e-6)
    i = s.integrator
    assert (i.max_bounces, i.min_bounces, i.diffuse_bounces, i.glossy_bounces, i.transmission_bounces,
            i.volume_bounces, i.transparent_min_bounces, i.transparent_max_bounces) == (11, 2, 3, 4, 5, 6, 1, 9)
    for value, expected in ((i.sample_clamp_direct, 1.25), (i.sample_clamp_indirect, 2.5),
                            (i.filter_glossy, 0.75), (i.film_exposure, 0.75),
                            (i.light_sampling_threshold, 0.125)):
        assert value == pytest.approx(expected, abs=1e-6)
    assert not i.reflective_caustics and i.refractive_caustics
    assert not i.use_light_tree
    assert i.direct_light_sampling == abi.DIRECT_LIGHT_NEE
    snap = imported.snapshot
    assert snap.world_visibility == abi.VIS_CAMERA | abi.VIS_GLOSSY | abi.VIS_SHADOW
    (light,) = snap.lights
    assert light.type == abi.LIGHT_AREA and light.flags & abi.LIGHT_FLAG_ELLIPSE
    assert light.size == pytest.approx(1.3, abs=1e-6) and light.size_y == pytest.approx(1.3, abs=1e-6)
    assert not light.flags & abi.LIGHT_FLAG_USE_MIS and not light.flags & abi.LIGHT_FLAG_CAST_SHADOW
    assert light.max_bounces == 13
    assert light.visibility_mask == abi.VIS_DIFFUSE | abi.VIS_TRANSMISSION | abi.VIS_SHADOW
    # material 0 is the default surface, authored materials follow in order
    assert imported.material_names == ["__cycles_default_surface__", "Raw Emissive"]
    m = snap.materials[1]
    assert m.emission_sampling == abi.EMISSION_SAMPLING_BACK
    assert [ins[0] for ins in m.instructions] == [abi.CLOSURE_EMISSION]
    assert np.allclose(m.parameters, [1.0, 0.5, 0.25, 3.0])
    cam = snap.camera
    assert cam.sensor_fit == 2 and cam.projection == 0
    assert cam.field_of_view == pytest.approx(0.6) and cam.horizontal_field_of_view == pytest.approx(0.9)
    assert (cam.near_clip, cam.far_clip) == (0.01, 100.0)


def test_legacy_seed_falls_back_to_authored_seed(tmp_path):
    scene = reference_scene()
    del scene["render"]["cycles"]["effective_seed"]
    imported = load_blender_scene_bundle(write(tmp_path / "b", scene))
    assert imported.ok() and imported.seed == 305419896


def test_defaults_when_keys_are_missing(tmp_path):
    scene = reference_scene()
    scene["render"] = {}
    scene["lights"][0] = {"type": "SPOT", "shape": "RECTANGLE", "size": 0.5}
    del scene["world"]["visibility"]
    imported = load_blender_scene_bundle(write(tmp_path / "b", scene))
    assert imported.ok(), errors(imported)
    assert (imported.width, imported.height, imported.samples) == (320, 240, 64)
    assert imported.settings.pixel_filter == abi.FILTER_BOX and imported.settings.filter_width == 1.0
    assert imported.settings.integrator.direct_light_sampling == abi.DIRECT_LIGHT_MIS
    (light,) = imported.snapshot.lights
    assert light.type == abi.LIGHT_SPOT and light.power == 1.0 and light.size_y == pytest.approx(0.5)
    assert light.spot_angle == pytest.approx(0.78539816339) and light.spot_smooth == pytest.approx(0.15)
    assert light.spread == pytest.approx(math.pi, abs=1e-6) and light.max_bounces == 1024
    assert light.flags == (abi.LIGHT_FLAG_NORMALIZE | abi.LIGHT_FLAG_SPHERE | abi.LIGHT_FLAG_USE_MIS |
                           abi.LIGHT_FLAG_CAST_SHADOW)
    assert light.visibility_mask == abi.VIS_ALL and imported.snapshot.world_visibility == abi.VIS_ALL
    # BOX is one whole pixel whatever filter_width says
    scene["render"] = {"pixel_filter_type": "BOX", "filter_width": 3.0}
    assert load_blender_scene_bundle(write(tmp_path / "c", scene)).settings.filter_width == 1.0


def test_visibility_mask_missing_keys_are_off():
    assert visibility_mask(None) == abi.VIS_ALL
    assert visibility_mask({}) == 0
    assert visibility_mask({"camera": True, "shadow": True}) == abi.VIS_CAMERA | abi.VIS_SHADOW


def test_light_power_exposure_and_temperature(tmp_path):
    scene = reference_scene()
    scene["lights"][0].update(type="POINT", energy=3.0, exposure=2.0, color=[0.5, 1.0, 1.0],
                              temperature_color=[1.0, 0.5, 0.25], use_soft_falloff=True)
    (light,) = load_blender_scene_bundle(write(tmp_path / "b", scene)).snapshot.lights
    assert light.type == abi.LIGHT_POINT and light.power == pytest.approx(12.0)
    assert list(light.color) == pytest.approx([0.5, 0.5, 0.25])
    assert not light.flags & abi.LIGHT_FLAG_SPHERE


@pytest.mark.parametrize("mutate, message", [
    (lambda s: s.update(schema="psycles.blender-scene.v9"), "unsupported Blender scene bundle schema"),
    (lambda s: s.pop("camera"), "bundle contains no active camera"),
    (lambda s: s["render"].update(pixel_filter_type="MITCHELL"), "unsupported Cycles pixel filter"),
    (lambda s: s["render"]["cycles"].update(direct_light_sampling_type="X"),
     "unsupported Cycles direct light sampling type"),
    (lambda s: s["materials"][0].update(emission_sampling="SIDEWAYS"),
     "unsupported Cycles emission sampling method"),
    (lambda s: s["camera"].update(type="PANO"), "panoramic cameras are not supported"),
    (lambda s: s["camera"]["dof"].update(enabled=True), "depth of field is not supported"),
    (lambda s: s["instances"].append({"geometry": 0, "geometry_type": "POINTCLOUD"}),
     "instance has unsupported geometry type"),
])
def test_rejections(tmp_path, mutate, message):
    scene = reference_scene()
    mutate(scene)
    imported = load_blender_scene_bundle(write(tmp_path / "b", scene))
    assert not imported.ok()
    assert any(message in m for m in errors(imported)), errors(imported)


def test_geometry_header_and_sections(tmp_path):
    imported = load_blender_scene_bundle(write(tmp_path / "a", reference_scene(), b"PSYGEO2\0"))
    assert "geometry.bin has an invalid header" in errors(imported)
    scene = reference_scene()
    scene["geometries"] = [{"name": "tri", "vertex_count": 3, "triangle_count": 1,
                            "positions": {"offset": 8}, "normals": {"offset": 8}, "uv": {"offset": 8},
                            "generated": {"offset": 8}, "indices": {"offset": 44},
                            "triangle_material_slots": {"offset": 56},
                            "triangle_random_per_island": {"offset": 56}, "material_slots": [None]}]
    imported = load_blender_scene_bundle(write(tmp_path / "b", scene, b"PSYGEO1\0" + bytes(40)))
    assert "failed to read geometry.bin section" in errors(imported)
    geometry = b"PSYGEO1\0" + np.zeros(9, np.float32).tobytes() + np.array([0, 1, 2, 0], np.uint32).tobytes()
    imported = load_blender_scene_bundle(write(tmp_path / "c", scene, geometry))
    assert imported.ok(), errors(imported)
    assert imported.snapshot.meshes[0].material_slots == [0]
    scene["schema"] = "psycles.blender-scene.v2"
    scene["geometries"][0].update(point_count=3, corner_count=4)
    imported = load_blender_scene_bundle(write(tmp_path / "d", scene, b"PSYGEO2\0" + geometry[8:]))
    assert "compact geometry has an invalid corner count" in errors(imported)


def test_unsupported_nodes_are_errors(tmp_path):
    scene = reference_scene()
    tree = scene["materials"][0]["node_tree"]
    tree["nodes"].append(bb.node("Noise", "TEX_NOISE"))
    tree["links"].append({"from_node": "Noise", "from_socket": "Color", "to_node": "Emission",
                          "to_socket": "Color"})
    imported = load_blender_scene_bundle(write(tmp_path / "a", scene))
    assert any("is linked" in m for m in errors(imported))
    scene = reference_scene()
    scene["materials"][0]["node_tree"]["nodes"][0]["type"] = "BSDF_PRINCIPLED"
    imported = load_blender_scene_bundle(write(tmp_path / "b", scene))
    assert any("unsupported shader node 'Emission' of type 'BSDF_PRINCIPLED'" in m for m in errors(imported))
    scene = reference_scene()
    scene["materials"][0]["node_tree"]["nodes"][0]["mute"] = True
    assert any("muted" in m for m in errors(load_blender_scene_bundle(write(tmp_path / "c", scene))))


def test_volume_material_is_reported(tmp_path):
    scene = reference_scene()
    tree = scene["materials"][0]["node_tree"]
    tree["surface_root"] = None
    tree["volume_root"] = {"node": "Add Volume", "socket": "Shader"}
    imported = load_blender_scene_bundle(write(tmp_path / "b", scene))
    assert any("volume shaders are not supported" in m for m in errors(imported))


def test_world_background_node(tmp_path):
    scene = reference_scene()
    scene["world"]["node_tree"] = {
        "nodes": [bb.node("Background", "BACKGROUND", [bb.socket("Color", [0.2, 0.4, 0.8, 1.0]),
                                                       bb.socket("Strength", 0.5)])],
        "links": [], "surface_root": {"node": "Background", "socket": "Background"}}
    imported = load_blender_scene_bundle(write(tmp_path / "b", scene))
    assert imported.ok(), errors(imported)
    assert imported.snapshot.world_color == pytest.approx((0.1, 0.2, 0.4))
    del scene["world"]["node_tree"]
    assert load_blender_scene_bundle(write(tmp_path / "c", scene)).snapshot.world_color == \
        pytest.approx((0.05, 0.05, 0.05))


def render_pair(snap, loaded, settings, samples=4):
    a, _ = ob.OracleScene(snap).render(settings, 0, samples, samples)
    b, _ = ob.OracleScene(loaded).render(settings, 0, samples, samples)
    for name in ("combined", "normal", "albedo"):
        assert np.array_equal(getattr(a, name).view(np.uint32), getattr(b, name).view(np.uint32)), name
    return a


@pytest.mark.parametrize("compact", [False, True], ids=["v1", "v2"])
def test_cornell_lamps_bundle_renders_like_the_snapshot(tmp_path, compact):
    snap = scenes.cornell_lamps()
    area = scenes.lamp(abi.LIGHT_AREA, scenes.translation(0.0, 0.9, 0.0) @ scenes.frame_z((0.0, -1.0, 0.0)),
                       (1.0, 0.8, 0.6), 6.0)
    area.size, area.size_y = 0.4, 0.25
    snap.lights.append(area)
    s = scenes.cornell_settings(48, 30)
    bb.snapshot_bundle(tmp_path / "b", snap, s, samples=4, compact=compact)
    imported = load_blender_scene_bundle(tmp_path / "b")
    assert imported.ok(), errors(imported)
    assert imported.samples == 4 and imported.seed == s.seed
    film = render_pair(snap, imported.snapshot, imported.settings)
    assert film.combined[..., :3].mean() > 0.01


@pytest.mark.parametrize("compact", [False, True], ids=["v1", "v2"])
def test_smooth_mixed_bundle_renders_like_the_snapshot(tmp_path, compact):
    snap = bb.smooth_sphere_snapshot()
    s = abi.default_settings(40, 30)
    s.seed = 7
    bb.snapshot_bundle(tmp_path / "b", snap, s, samples=4, compact=compact)
    imported = load_blender_scene_bundle(tmp_path / "b")
    assert imported.ok(), errors(imported)
    loaded = imported.snapshot
    if compact:
        # corner normals: one vertex per corner, same triangles
        assert loaded.meshes[0].positions.shape[0] == 3 * snap.meshes[0].triangles.shape[0]
    render_pair(snap, loaded, imported.settings)


def test_default_surface_warning(tmp_path):
    snap = scenes.cornell_box()
    snap.meshes[1] = copy.copy(snap.meshes[1])
    bb.snapshot_bundle(tmp_path / "b", snap, scenes.cornell_settings(8, 8))
    scene = json.load(open(tmp_path / "b" / "scene.json"))
    scene["geometries"][1]["material_slots"] = [None]
    with open(tmp_path / "b" / "scene.json", "w") as f:
        json.dump(scene, f)
    imported = load_blender_scene_bundle(tmp_path / "b")
    assert imported.ok()
    assert imported.snapshot.meshes[1].material_slots == [0]
    assert any("default surface" in d.message for d in imported.diagnostics if d.severity == "warning")
