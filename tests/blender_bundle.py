"""Writes Blender scene bundles for the loader tests.

The layout is the one the reference's exporter writes
(tools/export_psycles_scene.py): ``scene.json`` with column-major
transforms, node trees in Blender's vocabulary and byte offsets into
``geometry.bin``; v1 stores point-domain normals, v2 ("compact") names each
attribute's domain.  `snapshot_bundle` turns a SceneSnapshot into such a
bundle, so a loaded bundle can be rendered against the scene it came from.
"""
import json
import math
import os

import numpy as np

from psycles_amd import abi

_OP_NODE = {abi.CLOSURE_DIFFUSE: ("BSDF_DIFFUSE", "BSDF"), abi.CLOSURE_EMISSION: ("EMISSION", "Emission"),
            abi.CLOSURE_TRANSPARENT: ("BSDF_TRANSPARENT", "BSDF"), abi.CLOSURE_ADD: ("ADD_SHADER", "Shader"),
            abi.CLOSURE_MIX: ("MIX_SHADER", "Shader")}
_SAMPLING = {abi.EMISSION_SAMPLING_NONE: "NONE", abi.EMISSION_SAMPLING_AUTOMATIC: "AUTO",
             abi.EMISSION_SAMPLING_FRONT: "FRONT", abi.EMISSION_SAMPLING_BACK: "BACK",
             abi.EMISSION_SAMPLING_FRONT_BACK: "FRONT_BACK"}


def socket(identifier, value, kind="NodeSocketFloat"):
    return {"identifier": identifier, "name": identifier, "type": kind, "linked": False,
            "default": value}


def node(name, kind, inputs=()):
    return {"name": name, "label": "", "type": kind, "bl_idname": "", "inputs": list(inputs),
            "outputs": [], "properties": {}, "special": {}, "image": None, "node_tree": None}


def node_tree(material):
    """Blender node tree for a closure program."""
    p = [float(v) for v in material.parameters]
    nodes, links = [], []
    for i, (op, a, b, factor, color, strength, roughness) in enumerate(material.instructions):
        name = f"n{i}"
        if op == abi.CLOSURE_DIFFUSE:
            nodes.append(node(name, "BSDF_DIFFUSE", [socket("Color", p[color:color + 3] + [1.0], "NodeSocketColor"),
                                                     socket("Roughness", p[roughness])]))
        elif op == abi.CLOSURE_EMISSION:
            nodes.append(node(name, "EMISSION", [socket("Color", p[color:color + 3] + [1.0], "NodeSocketColor"),
                                                 socket("Strength", p[strength])]))
        elif op == abi.CLOSURE_TRANSPARENT:
            nodes.append(node(name, "BSDF_TRANSPARENT",
                              [socket("Color", p[color:color + 3] + [1.0], "NodeSocketColor")]))
        elif op in (abi.CLOSURE_ADD, abi.CLOSURE_MIX):
            inputs = [socket("Fac", p[factor])] if op == abi.CLOSURE_MIX else []
            nodes.append(node(name, _OP_NODE[op][0], inputs))
            for child, target in ((a, "Shader"), (b, "Shader_001")):
                links.append({"from_node": f"n{child}", "from_socket": _OP_NODE[material.instructions[child][0]][1],
                              "to_node": name, "to_socket": target})
        else:
            raise ValueError(op)
    root = material.root
    return {"name": material.name, "nodes": nodes, "links": links,
            "surface_root": {"node": f"n{root}", "socket": _OP_NODE[material.instructions[root][0]][1]},
            "volume_root": None, "displacement_root": None}


def column_major(m):
    return [float(v) for v in np.asarray(m, dtype=np.float32).T.reshape(-1)]


def visibility(mask):
    names = ("camera", "diffuse", "glossy", "transmission", "shadow", "volume_scatter")
    return {n: bool(mask & (1 << i)) for i, n in enumerate(names)}


_LIGHT_TYPE = {abi.LIGHT_POINT: "POINT", abi.LIGHT_SPOT: "SPOT", abi.LIGHT_AREA: "AREA", abi.LIGHT_DISTANT: "SUN"}


def light_json(l):
    transform = np.asarray(list(l.transform), dtype=np.float32).reshape(4, 4).T   # stored column-major
    kind = _LIGHT_TYPE[l.type]
    ellipse = bool(l.flags & abi.LIGHT_FLAG_ELLIPSE)
    return {"name": f"light{kind}", "type": kind, "transform": column_major(transform),
            "color": [float(l.color[i]) for i in range(3)], "energy": float(l.power),
            "shape": ("ELLIPSE" if ellipse else "RECTANGLE") if kind == "AREA" else "POINT",
            "size": float(l.size), "size_y": float(l.size_y), "spread": float(l.spread),
            "spot_size": float(l.spot_angle), "spot_blend": float(l.spot_smooth), "angle": float(l.angle),
            "normalize": bool(l.flags & abi.LIGHT_FLAG_NORMALIZE),
            "use_soft_falloff": not (l.flags & abi.LIGHT_FLAG_SPHERE),
            "use_multiple_importance_sampling": bool(l.flags & abi.LIGHT_FLAG_USE_MIS),
            "cast_shadow": bool(l.flags & abi.LIGHT_FLAG_CAST_SHADOW),
            "max_bounces": int(l.max_bounces), "visibility": visibility(l.visibility_mask), "node_tree": None}


class _Blob:
    def __init__(self, magic):
        self.data = bytearray(magic)

    def add(self, array):
        while len(self.data) % 16:
            self.data += b"\0"
        offset = len(self.data)
        self.data += np.ascontiguousarray(array).tobytes()
        return {"offset": offset}


def _mesh_json(mesh, blob, compact, material_names):
    t = np.asarray(mesh.triangles, dtype=np.uint32)
    v = np.asarray(mesh.positions, dtype=np.float32)
    count = t.shape[0]
    normals = mesh.normals if mesh.normals is not None else np.zeros_like(v)
    smooth = mesh.smooth if mesh.smooth is not None else np.zeros(count, np.uint8)
    slots = (mesh.triangle_material_slots if mesh.triangle_material_slots is not None
             else np.zeros(count, np.uint32))
    g = {"name": mesh.name, "triangle_count": count,
         "material_slots": [material_names[s] for s in mesh.material_slots]}
    g["positions"] = blob.add(v)
    if compact:
        # corner-domain normals: the point normals of each triangle corner
        g.update(point_count=v.shape[0], corner_count=3 * count, normal_domain="CORNER",
                 uv_domain="POINT", uv_tangent_domain="POINT", generated_domain="POINT")
        g["normals"] = blob.add(np.asarray(normals, np.float32)[t.reshape(-1)])
    else:
        g["vertex_count"] = v.shape[0]
        g["normals"] = blob.add(np.asarray(normals, np.float32))
    g["uv"] = blob.add(np.zeros((v.shape[0], 2), np.float32))
    g["generated"] = blob.add(v)
    g["indices"] = blob.add(t)
    g["triangle_material_slots"] = blob.add(np.asarray(slots, np.uint32))
    g["triangle_smooth"] = blob.add(np.asarray(smooth, np.uint32))
    g["triangle_random_per_island"] = blob.add(np.zeros(count, np.float32))
    return g


def render_json(settings, samples):
    i = settings.integrator
    filters = {abi.FILTER_BOX: "BOX", abi.FILTER_GAUSSIAN: "GAUSSIAN", abi.FILTER_BLACKMAN_HARRIS: "BLACKMAN_HARRIS"}
    modes = {abi.DIRECT_LIGHT_MIS: "MULTIPLE_IMPORTANCE_SAMPLING", abi.DIRECT_LIGHT_FORWARD: "FORWARD_PATH_TRACING",
             abi.DIRECT_LIGHT_NEE: "NEXT_EVENT_ESTIMATION"}
    return {"width": int(settings.width), "height": int(settings.height), "percentage": 100,
            "transparent": bool(settings.transparent_background),
            "pixel_filter_type": filters[settings.pixel_filter], "filter_width": float(settings.filter_width),
            "pass_alpha_threshold": float(settings.pass_alpha_threshold),
            "cycles": {"samples": samples, "seed": int(settings.seed), "effective_seed": int(settings.seed),
                       "max_bounces": i.max_bounces, "min_light_bounces": i.min_bounces,
                       "diffuse_bounces": i.diffuse_bounces, "glossy_bounces": i.glossy_bounces,
                       "transmission_bounces": i.transmission_bounces, "volume_bounces": i.volume_bounces,
                       "min_transparent_bounces": i.transparent_min_bounces,
                       "transparent_max_bounces": i.transparent_max_bounces,
                       "sample_clamp_direct": i.sample_clamp_direct,
                       "sample_clamp_indirect": i.sample_clamp_indirect, "blur_glossy": i.filter_glossy,
                       "film_exposure": i.film_exposure, "light_sampling_threshold": i.light_sampling_threshold,
                       "caustics_reflective": bool(i.reflective_caustics),
                       "caustics_refractive": bool(i.refractive_caustics),
                       "use_light_tree": bool(i.use_light_tree),
                       "direct_light_sampling_type": modes[i.direct_light_sampling]}}


def snapshot_bundle(directory, snapshot, settings, samples=16, compact=False):
    """Write `snapshot` + `settings` as a bundle in `directory`."""
    os.makedirs(directory, exist_ok=True)
    blob = _Blob(b"PSYGEO2\0" if compact else b"PSYGEO1\0")
    names = [m.name or f"material{i}" for i, m in enumerate(snapshot.materials)]
    materials = [{"name": names[i], "emission_sampling": _SAMPLING[m.emission_sampling],
                  "node_tree": node_tree(m)} for i, m in enumerate(snapshot.materials)]
    geometries = [_mesh_json(m, blob, compact, names) for m in snapshot.meshes]
    instances = [{"name": inst.name, "geometry": inst.mesh, "geometry_type": "MESH",
                  "transform": column_major(inst.transform), "visibility": visibility(inst.visibility),
                  "shadow_terminator_geometry_offset": inst.shadow_terminator_geometry_offset}
                 for inst in snapshot.instances]
    cam = snapshot.camera
    fit = {0: "HORIZONTAL", 1: "VERTICAL", 2: "AUTO"}[cam.sensor_fit]
    hfov = cam.field_of_view if cam.horizontal_field_of_view is None else cam.horizontal_field_of_view
    camera = {"name": "Camera", "type": "ORTHO" if cam.projection == 1 else "PERSP",
              "transform": column_major(cam.transform), "angle": cam.field_of_view,
              "angle_x": hfov, "angle_y": cam.field_of_view, "sensor_fit": fit,
              "ortho_scale": cam.orthographic_scale, "shift_x": 0.0, "shift_y": 0.0,
              "clip_start": cam.near_clip, "clip_end": cam.far_clip, "lens": 50.0,
              "dof": {"enabled": False, "focus_distance": 10.0, "fstop": 2.8}}
    world = {"name": "World", "color": [float(c) for c in snapshot.world_color],
             "visibility": visibility(snapshot.world_visibility), "node_tree": None}
    scene = {"schema": "psycles.blender-scene.v2" if compact else "psycles.blender-scene.v1",
             "images": [], "node_groups": [], "materials": materials, "geometries": geometries,
             "instances": instances, "lights": [light_json(l) for l in snapshot.lights],
             "world": world, "world_environment": None, "render": render_json(settings, samples),
             "camera": camera}
    with open(os.path.join(directory, "scene.json"), "w") as f:
        json.dump(scene, f, indent=1)
    with open(os.path.join(directory, "geometry.bin"), "wb") as f:
        f.write(bytes(blob.data))
    return directory


def smooth_sphere_snapshot():
    """Two smooth-shaded icospheres (per-vertex normals) on a diffuse floor
    under an area lamp, with a mix/add/transparent material on one sphere."""
    from psycles_amd import scenes
    p, t = scenes.geodesic_icosphere(4, displacement=0.0)
    n = p / np.linalg.norm(p, axis=1, keepdims=True)
    b = scenes.ProgramBuilder()
    glassy = b.build(b.add(b.mix(b.diffuse((0.7, 0.3, 0.2)), b.transparent((0.9, 0.9, 1.0)), 0.35),
                           b.emission((0.2, 0.4, 1.0), 0.5)), name="mixed")
    materials = [scenes.diffuse_material((0.6, 0.6, 0.6), name="floor"),
                 scenes.diffuse_material((0.2, 0.5, 0.8), 0.5, name="blue"), glassy]
    normals = n.astype(np.float32)
    smooth = np.ones(t.shape[0], np.uint8)
    sphere = scenes.MeshData(p.astype(np.float32), t, [1], normals=normals, smooth=smooth, name="sphere")
    sphere_mixed = scenes.MeshData(p.astype(np.float32), t, [2], normals=normals, smooth=smooth,
                                   name="sphere_mixed")
    fp, ft = scenes.quad([[-3, -1, -3], [-3, -1, 3], [3, -1, 3], [3, -1, -3]])
    floor = scenes.MeshData(fp, ft, [0], name="floor")
    instances = [scenes.InstanceData(0, scenes.translation(-0.7, 0.0, 0.0) @ scenes.scale(0.8, 0.8, 0.8)),
                 scenes.InstanceData(2, scenes.translation(0.8, -0.2, 0.3) @ scenes.scale(0.6, 0.6, 0.6)),
                 scenes.InstanceData(1, scenes.identity())]
    area = scenes.lamp(abi.LIGHT_AREA, scenes.translation(0.0, 2.5, 1.0) @ scenes.frame_z((0.0, -1.0, -0.3)),
                       (1.0, 0.95, 0.9), 40.0)
    area.size, area.size_y = 1.2, 0.6
    area.flags |= abi.LIGHT_FLAG_ELLIPSE
    camera = scenes.CameraData(scenes.translation(0.0, 0.3, 4.0), math.radians(45.0), 1.0e-3, 100.0)
    return scenes.SceneSnapshot(materials, [sphere, floor, sphere_mixed], instances, camera,
                                (0.05, 0.06, 0.08), lights=[area], name="smooth_spheres")
