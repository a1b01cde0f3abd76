"""Blender scene bundle loader for the closure subset this renderer draws.

A bundle is a directory holding ``scene.json`` (schema
``psycles.blender-scene.v1`` or ``.v2``) and ``geometry.bin`` (magic
``PSYGEO1\\0`` or ``PSYGEO2\\0``), as written by the reference's Blender
exporter (tools/export_psycles_scene.py).  `load_blender_scene_bundle`
follows src/adapter/blender_scene.cpp's mapping field for field:

* render size, percentage, samples, seed (``effective_seed`` before
  ``seed``), pixel filter, transparency and every Cycles integrator key
  (blender_scene.cpp:476-693);
* the active camera (blender_scene.cpp:694-768);
* mesh geometries from byte offsets into geometry.bin
  (blender_scene.cpp:770-1163), material slots by name with the default
  surface in slot 0 of the material list (blender_scene.cpp:385-393);
* instances (blender_scene.cpp:1280-1358), lights (1360-1491) and a
  constant world (1492-1531);
* ray visibility dictionaries (src/adapter/blender_scene_json.cpp:91-120).

Shader node trees are lowered to closure programs for the node types the
device evaluates - BSDF_DIFFUSE, EMISSION, BSDF_TRANSPARENT, ADD_SHADER,
MIX_SHADER, BACKGROUND for the world, REROUTE - with unlinked value inputs,
following src/adapter/blender_graph_lower_closures.cpp:211-540.  Anything
else (textures, Principled, volumes, curves, light shaders, panoramic or
depth-of-field cameras) becomes an error diagnostic and the import is not
``ok()``: the scene is never silently approximated, with one stated
exception - Cycles' default surface is a Principled BSDF of base colour 0.8
(blender_scene_json.cpp:330-342), which this subset draws as a Diffuse BSDF
of the same colour and reports with a warning when a mesh uses it.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .scenes import CameraData, InstanceData, Material, MeshData, ProgramBuilder, SceneSnapshot

SCHEMA_V1 = "psycles.blender-scene.v1"
SCHEMA_V2 = "psycles.blender-scene.v2"
MAGIC_V1 = b"PSYGEO1\0"
MAGIC_V2 = b"PSYGEO2\0"

_VISIBILITY_KEYS = (("camera", abi.VIS_CAMERA), ("diffuse", abi.VIS_DIFFUSE),
                    ("glossy", abi.VIS_GLOSSY), ("transmission", abi.VIS_TRANSMISSION),
                    ("shadow", abi.VIS_SHADOW), ("volume_scatter", abi.VIS_VOLUME_SCATTER))

_EMISSION_SAMPLING = {"NONE": abi.EMISSION_SAMPLING_NONE, "AUTO": abi.EMISSION_SAMPLING_AUTOMATIC,
                      "FRONT": abi.EMISSION_SAMPLING_FRONT, "BACK": abi.EMISSION_SAMPLING_BACK,
                      "FRONT_BACK": abi.EMISSION_SAMPLING_FRONT_BACK}

_DIRECT_LIGHT = {"MULTIPLE_IMPORTANCE_SAMPLING": abi.DIRECT_LIGHT_MIS,
                 "FORWARD_PATH_TRACING": abi.DIRECT_LIGHT_FORWARD,
                 "NEXT_EVENT_ESTIMATION": abi.DIRECT_LIGHT_NEE}

DEFAULT_SURFACE_COLOR = (0.8, 0.8, 0.8)


class BundleError(Exception):
    """A malformed bundle: the import stops (the reference throws)."""


@dataclass
class Diagnostic:
    severity: str        # "error" or "warning"
    message: str


@dataclass
class BlenderImport:
    """ImportedBlenderScene counterpart (include/psycles/adapter/blender_scene.h)."""

    snapshot: SceneSnapshot | None = None
    settings: abi.RenderSettings | None = None
    samples: int = 0
    adaptive_sampling: bool = False
    denoising: bool = False
    material_names: list = field(default_factory=list)
    diagnostics: list = field(default_factory=list)

    def ok(self) -> bool:
        return self.snapshot is not None and not any(d.severity == "error" for d in self.diagnostics)

    @property
    def width(self) -> int:
        return int(self.settings.width) if self.settings is not None else 0

    @property
    def height(self) -> int:
        return int(self.settings.height) if self.settings is not None else 0

    @property
    def seed(self) -> int:
        return int(self.settings.seed) if self.settings is not None else 0


# --------------------------------------------------------------------------
# JSON helpers with the reference's defaults (blender_scene_json.cpp)
# --------------------------------------------------------------------------

def _member(obj, name):
    return obj.get(name) if isinstance(obj, dict) else None


def _number(value, default=0.0) -> float:
    return float(value) if isinstance(value, (int, float)) and not isinstance(value, bool) else default


def _unsigned(value, default=0) -> int:
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        return default
    return int(value) if value >= 0 else default


def _boolean(value, default=False) -> bool:
    return value if isinstance(value, bool) else default


def _text(value, default="") -> str:
    return value if isinstance(value, str) else default


def _float3(value, default) -> tuple:
    if isinstance(value, list) and len(value) >= 3:
        return tuple(_number(value[i], default[i]) for i in range(3))
    return tuple(default)


def visibility_mask(value) -> int:
    """Bits of the visibility dictionary; a missing dictionary means all."""
    mask = 0
    for name, bit in _VISIBILITY_KEYS:
        if _boolean(_member(value, name), value is None):
            mask |= bit
    return mask


def column_major_matrix(value) -> np.ndarray:
    """16 column-major numbers (the exporter's _column_major) -> row-major
    4x4; anything else is the zero matrix, as Mat4f{} is."""
    if not isinstance(value, list) or len(value) != 16:
        return np.zeros((4, 4), dtype=np.float32)
    return np.asarray([_number(v) for v in value], dtype=np.float32).reshape(4, 4).T.copy()


# --------------------------------------------------------------------------
# Shader node trees -> closure programs
# --------------------------------------------------------------------------

class _Lowering:
    """One material's node tree lowered into a ProgramBuilder."""

    def __init__(self, owner: str, tree: dict, diagnostics: list):
        self.owner = owner
        self.nodes = {n.get("name"): n for n in tree.get("nodes") or [] if isinstance(n, dict)}
        self.links = {}
        for link in tree.get("links") or []:
            if isinstance(link, dict):
                self.links[(link.get("to_node"), link.get("to_socket"))] = (
                    link.get("from_node"), link.get("from_socket"))
        self.diagnostics = diagnostics
        self.builder = ProgramBuilder()
        self.failed = False
        self.visiting: set = set()

    def error(self, message: str) -> int:
        self.failed = True
        self.diagnostics.append(Diagnostic("error", f"{self.owner}: {message}"))
        return self.null()

    def null(self) -> int:
        return self.builder._emit(abi.CLOSURE_NULL)

    def _input(self, node: dict, identifier: str):
        for socket in node.get("inputs") or []:
            if socket.get("identifier") == identifier:
                return socket
        return None

    def value(self, node: dict, identifier: str, default):
        """Unlinked default of a value input; a link is an unsupported
        texture or math graph in this subset."""
        if (node.get("name"), identifier) in self.links:
            self.error(f"node '{node.get('name')}' input '{identifier}' is linked; "
                       "only constant inputs are supported")
            return default
        socket = self._input(node, identifier)
        raw = socket.get("default") if socket is not None else None
        if isinstance(default, tuple):
            return _float3(raw, default)
        return _number(raw, default)

    def closure_input(self, node: dict, identifier: str) -> int:
        source = self.links.get((node.get("name"), identifier))
        if source is None:
            return self.null()
        return self.lower(source[0])

    def lower(self, name) -> int:
        node = self.nodes.get(name)
        if node is None:
            return self.error(f"link from missing node '{name}'")
        if name in self.visiting:
            return self.error(f"node '{name}' is part of a cycle")
        self.visiting.add(name)
        try:
            return self._lower(node)
        finally:
            self.visiting.discard(name)

    def _lower(self, node: dict) -> int:
        kind = node.get("type")
        b = self.builder
        if _boolean(node.get("mute")):
            return self.error(f"node '{node.get('name')}' is muted; muted nodes are not supported")
        if kind == "REROUTE":
            return self.closure_input(node, "Input")
        if kind == "BSDF_DIFFUSE":
            if (node.get("name"), "Normal") in self.links:
                return self.error(f"node '{node.get('name')}' has a linked Normal")
            return b.diffuse(self.value(node, "Color", DEFAULT_SURFACE_COLOR),
                             self.value(node, "Roughness", 0.0))
        if kind in ("EMISSION", "BACKGROUND"):
            return b.emission(self.value(node, "Color", (1.0, 1.0, 1.0)),
                              self.value(node, "Strength", 1.0))
        if kind == "BSDF_TRANSPARENT":
            return b.transparent(self.value(node, "Color", (1.0, 1.0, 1.0)))
        if kind == "ADD_SHADER":
            return b.add(self.closure_input(node, "Shader"), self.closure_input(node, "Shader_001"))
        if kind == "MIX_SHADER":
            factor = self.value(node, "Fac", 0.5)
            return b.mix(self.closure_input(node, "Shader"), self.closure_input(node, "Shader_001"),
                         factor)
        return self.error(f"unsupported shader node '{node.get('name')}' of type '{kind}'")


def lower_surface(owner: str, holder: dict, diagnostics: list,
                  emission_sampling=abi.EMISSION_SAMPLING_AUTOMATIC) -> Material:
    """Closure program of `holder`'s node_tree surface output.  A material
    without a connected surface has the empty (null) closure."""
    tree = _member(holder, "node_tree")
    lowering = _Lowering(owner, tree if isinstance(tree, dict) else {}, diagnostics)
    root_ref = _member(tree, "surface_root")
    if isinstance(tree, dict) and isinstance(tree.get("volume_root"), dict):
        lowering.error("volume shaders are not supported")
    if isinstance(tree, dict) and isinstance(tree.get("displacement_root"), dict):
        lowering.error("displacement is not supported")
    if isinstance(root_ref, dict):
        root = lowering.lower(root_ref.get("node"))
    else:
        root = lowering.null()
    return lowering.builder.build(root, emission_sampling, name=owner)


def _constant_emission(material: Material):
    """(colour, strength) of a program that is one emission closure."""
    if len(material.instructions) != 1 or material.instructions[0][0] != abi.CLOSURE_EMISSION:
        return None
    op, _a, _b, _f, color, strength, _r = material.instructions[0]
    p = material.parameters
    return tuple(float(v) for v in p[color:color + 3]), float(p[strength])


# --------------------------------------------------------------------------
# geometry.bin
# --------------------------------------------------------------------------

class _GeometryFile:
    def __init__(self, path: str):
        try:
            with open(path, "rb") as f:
                self.data = f.read()
        except OSError as exc:
            raise BundleError("failed to open geometry.bin") from exc

    def values(self, geometry: dict, section: str, count: int, dtype) -> np.ndarray:
        offset = _unsigned(_member(_member(geometry, section), "offset"))
        size = count * np.dtype(dtype).itemsize
        if offset + size > len(self.data):
            raise BundleError("failed to read geometry.bin section")
        return np.frombuffer(self.data, dtype=dtype, count=count, offset=offset).copy()


def _domain(value) -> str:
    name = _text(value)
    if name in ("", "POINT"):
        return "point"
    if name == "CORNER":
        return "corner"
    if name == "FACE":
        return "face"
    raise BundleError(f"unsupported mesh attribute domain: {name}")


def _mesh(geometry: dict, geo: _GeometryFile, compact: bool, material_ids: dict) -> MeshData:
    triangles = _unsigned(_member(geometry, "triangle_count"))
    points = _unsigned(_member(geometry, "point_count" if compact else "vertex_count"))
    corners = triangles * 3
    if compact and _unsigned(_member(geometry, "corner_count")) != corners:
        raise BundleError("compact geometry has an invalid corner count")
    normal_domain = _domain(_member(geometry, "normal_domain")) if compact else "point"
    count = {"point": points, "corner": corners, "face": triangles}
    positions = geo.values(geometry, "positions", points * 3, np.float32).reshape(points, 3)
    normals = geo.values(geometry, "normals", count[normal_domain] * 3, np.float32).reshape(-1, 3)
    indices = geo.values(geometry, "indices", corners, np.uint32).reshape(triangles, 3)
    slots = geo.values(geometry, "triangle_material_slots", triangles, np.uint32)
    smooth = (geo.values(geometry, "triangle_smooth", triangles, np.uint32)
              if _member(geometry, "triangle_smooth") is not None else np.zeros(triangles, np.uint32))
    if triangles and int(indices.max()) >= points:
        raise BundleError("mesh index out of range")
    if normal_domain != "point":
        # Corner and face normals: one vertex per corner, so each triangle
        # carries its own normals; the triangle geometry is unchanged.
        positions = positions[indices.reshape(-1)]
        normals = normals if normal_domain == "corner" else np.repeat(normals, 3, axis=0)
        indices = np.arange(corners, dtype=np.uint32).reshape(triangles, 3)
    material_slots = []
    for slot in _member(geometry, "material_slots") or []:
        material_slots.append(material_ids.get(slot, 0) if isinstance(slot, str) else 0)
    if not material_slots:
        material_slots = [0]
    return MeshData(positions=np.ascontiguousarray(positions, dtype=np.float32),
                    triangles=np.ascontiguousarray(indices, dtype=np.uint32),
                    material_slots=material_slots,
                    triangle_material_slots=slots,
                    normals=np.ascontiguousarray(normals, dtype=np.float32),
                    smooth=(smooth != 0).astype(np.uint8),
                    name=_text(_member(geometry, "name")))


# --------------------------------------------------------------------------
# lights, camera, render settings
# --------------------------------------------------------------------------

_LIGHT_TYPES = {"AREA": abi.LIGHT_AREA, "SUN": abi.LIGHT_DISTANT, "SPOT": abi.LIGHT_SPOT}


def _light(light: dict) -> abi.Light:
    color = _float3(_member(light, "color"), (1.0, 1.0, 1.0))
    temperature = _float3(_member(light, "temperature_color"), (1.0, 1.0, 1.0))
    size = _number(_member(light, "size"), 0.0)
    shape = _text(_member(light, "shape"), "POINT")
    size_y = size if shape in ("SQUARE", "DISK") else _number(_member(light, "size_y"), size)
    l = abi.Light()
    l.type = _LIGHT_TYPES.get(_text(_member(light, "type")), abi.LIGHT_POINT)
    l.visibility_mask = visibility_mask(_member(light, "visibility"))
    for i, v in enumerate(column_major_matrix(_member(light, "transform")).T.reshape(-1)):
        l.transform[i] = float(v)
    for i in range(3):
        l.color[i] = color[i] * temperature[i]
    l.power = _number(_member(light, "energy"), 1.0) * 2.0 ** _number(_member(light, "exposure"), 0.0)
    l.size = size
    l.size_y = size_y
    l.spread = _number(_member(light, "spread"), 3.14159265)
    l.spot_angle = _number(_member(light, "spot_size"), 0.78539816339)
    l.spot_smooth = _number(_member(light, "spot_blend"), 0.15)
    l.angle = _number(_member(light, "angle"), 0.0)
    flags = 0
    if _boolean(_member(light, "normalize"), True):
        flags |= abi.LIGHT_FLAG_NORMALIZE
    if shape in ("DISK", "ELLIPSE"):
        flags |= abi.LIGHT_FLAG_ELLIPSE
    if not _boolean(_member(light, "use_soft_falloff"), False):
        flags |= abi.LIGHT_FLAG_SPHERE
    if _boolean(_member(light, "use_multiple_importance_sampling"), True):
        flags |= abi.LIGHT_FLAG_USE_MIS
    if _boolean(_member(light, "cast_shadow"), True):
        flags |= abi.LIGHT_FLAG_CAST_SHADOW
    l.flags = flags
    l.max_bounces = _unsigned(_member(light, "max_bounces"), 1024)
    return l


_SENSOR_FIT = {"HORIZONTAL": 0, "VERTICAL": 1}


def _camera(camera: dict, diagnostics: list) -> CameraData:
    kind = _text(_member(camera, "type"))
    if kind == "PANO":
        diagnostics.append(Diagnostic("error", "panoramic cameras are not supported"))
    dof = _member(camera, "dof")
    if _boolean(_member(dof, "enabled")):
        diagnostics.append(Diagnostic("error", "depth of field is not supported"))
    if _number(_member(camera, "shift_x")) != 0.0 or _number(_member(camera, "shift_y")) != 0.0:
        diagnostics.append(Diagnostic("error", "lens shift is not supported"))
    angle = _number(_member(camera, "angle"), 0.7853982)
    return CameraData(transform=column_major_matrix(_member(camera, "transform")),
                      field_of_view=_number(_member(camera, "angle_y"), angle),
                      horizontal_field_of_view=_number(_member(camera, "angle_x"), angle),
                      near_clip=_number(_member(camera, "clip_start"), 1.0e-4),
                      far_clip=_number(_member(camera, "clip_end"), 1.0e5),
                      sensor_fit=_SENSOR_FIT.get(_text(_member(camera, "sensor_fit")), 2),
                      projection=1 if kind == "ORTHO" else 0,
                      orthographic_scale=_number(_member(camera, "ortho_scale"), 1.0))


def _render_settings(render: dict, result: BlenderImport) -> abi.RenderSettings:
    width = _unsigned(_member(render, "width"), 320)
    height = _unsigned(_member(render, "height"), 240)
    percentage = _unsigned(_member(render, "percentage"), 100)
    width = max(width * percentage // 100, 1)
    height = max(height * percentage // 100, 1)
    s = abi.default_settings(width, height)
    cycles = _member(render, "cycles")
    result.samples = _unsigned(_member(cycles, "samples"), 64)
    s.pass_alpha_threshold = min(max(_number(_member(render, "pass_alpha_threshold"), 0.5), 0.0), 1.0)
    s.seed = _unsigned(_member(cycles, "effective_seed"), _unsigned(_member(cycles, "seed"))) & 0xFFFFFFFF
    result.adaptive_sampling = _boolean(_member(cycles, "use_adaptive_sampling"))
    result.denoising = _boolean(_member(cycles, "use_denoising"))
    i = s.integrator
    for field_name, key in (("max_bounces", "max_bounces"), ("min_bounces", "min_light_bounces"),
                            ("diffuse_bounces", "diffuse_bounces"), ("glossy_bounces", "glossy_bounces"),
                            ("transmission_bounces", "transmission_bounces"),
                            ("volume_bounces", "volume_bounces"),
                            ("transparent_min_bounces", "min_transparent_bounces"),
                            ("transparent_max_bounces", "transparent_max_bounces")):
        setattr(i, field_name, _unsigned(_member(cycles, key), getattr(i, field_name)))
    for field_name, key in (("sample_clamp_direct", "sample_clamp_direct"),
                            ("sample_clamp_indirect", "sample_clamp_indirect"),
                            ("filter_glossy", "blur_glossy"), ("film_exposure", "film_exposure"),
                            ("light_sampling_threshold", "light_sampling_threshold")):
        setattr(i, field_name, max(_number(_member(cycles, key), getattr(i, field_name)), 0.0))
    i.reflective_caustics = int(_boolean(_member(cycles, "caustics_reflective"), bool(i.reflective_caustics)))
    i.refractive_caustics = int(_boolean(_member(cycles, "caustics_refractive"), bool(i.refractive_caustics)))
    i.use_light_tree = int(_boolean(_member(cycles, "use_light_tree"), bool(i.use_light_tree)))
    mode = _text(_member(cycles, "direct_light_sampling_type"), "MULTIPLE_IMPORTANCE_SAMPLING")
    if mode not in _DIRECT_LIGHT:
        raise BundleError(f"unsupported Cycles direct light sampling type: {mode}")
    i.direct_light_sampling = _DIRECT_LIGHT[mode]
    if result.adaptive_sampling:
        result.diagnostics.append(Diagnostic(
            "warning", "Cycles adaptive sampling is enabled in the source scene; a fixed sample "
                       "count is rendered"))
    if result.denoising:
        result.diagnostics.append(Diagnostic(
            "warning", "Cycles denoising is enabled in the source scene; passes are un-denoised"))
    s.transparent_background = int(_boolean(_member(render, "transparent")))
    requested_width = max(_number(_member(render, "filter_width"), 1.0), 1.0e-5)
    kind = _text(_member(render, "pixel_filter_type"))
    if kind in ("", "BOX"):
        s.pixel_filter, s.filter_width = abi.FILTER_BOX, 1.0
    elif kind == "GAUSSIAN":
        s.pixel_filter, s.filter_width = abi.FILTER_GAUSSIAN, requested_width
    elif kind == "BLACKMAN_HARRIS":
        s.pixel_filter, s.filter_width = abi.FILTER_BLACKMAN_HARRIS, requested_width
    else:
        raise BundleError(f"unsupported Cycles pixel filter: {kind}")
    return s


# --------------------------------------------------------------------------
# the bundle
# --------------------------------------------------------------------------

def load_blender_scene_bundle(directory) -> BlenderImport:
    """load_blender_scene_bundle (src/adapter/blender_scene.cpp:300-1560):
    a SceneSnapshot, RenderSettings and sample count for the bundle in
    `directory`, with diagnostics.  A malformed bundle returns an import
    whose ``snapshot`` is None and whose diagnostics say why."""
    result = BlenderImport()
    try:
        with open(os.path.join(directory, "scene.json"), "r", encoding="utf-8") as f:
            root = json.load(f)
    except (OSError, ValueError) as exc:
        result.diagnostics.append(Diagnostic("error", f"failed to read scene.json: {exc}"))
        return result
    schema = _text(_member(root, "schema"))
    if schema not in (SCHEMA_V1, SCHEMA_V2):
        result.diagnostics.append(Diagnostic("error", "unsupported Blender scene bundle schema"))
        return result
    compact = schema == SCHEMA_V2
    try:
        snapshot = _load(directory, root, compact, result)
    except BundleError as exc:
        result.diagnostics.append(Diagnostic("error", str(exc)))
        return result
    result.snapshot = snapshot
    return result


def _load(directory, root: dict, compact: bool, result: BlenderImport) -> SceneSnapshot:
    diagnostics = result.diagnostics
    b = ProgramBuilder()
    materials = [b.build(b.diffuse(DEFAULT_SURFACE_COLOR), name="__cycles_default_surface__")]
    material_ids: dict = {}
    for material in _member(root, "materials") or []:
        name = _text(_member(material, "name"))
        sampling = _text(_member(material, "emission_sampling"), "AUTO")
        if sampling not in _EMISSION_SAMPLING:
            raise BundleError(f"unsupported Cycles emission sampling method: {sampling}")
        material_ids[name] = len(materials)
        materials.append(lower_surface(name, material, diagnostics, _EMISSION_SAMPLING[sampling]))
    render = _member(root, "render")
    result.settings = _render_settings(render, result)
    camera = _member(root, "camera")
    if not isinstance(camera, dict):
        raise BundleError("bundle contains no active camera")
    camera_data = _camera(camera, diagnostics)
    geo = _GeometryFile(os.path.join(directory, "geometry.bin"))
    if geo.data[:8] != (MAGIC_V2 if compact else MAGIC_V1):
        raise BundleError("geometry.bin has an invalid header")
    meshes = [_mesh(g, geo, compact, material_ids) for g in _member(root, "geometries") or []]
    instances = []
    for instance in _member(root, "instances") or []:
        kind = _text(_member(instance, "geometry_type"), "MESH")
        if kind not in ("MESH", "CURVE"):
            raise BundleError(f"instance has unsupported geometry type '{kind}'")
        index = _unsigned(_member(instance, "geometry"))
        if kind == "CURVE":
            diagnostics.append(Diagnostic("error", "curve geometry is not supported"))
            continue
        if index >= len(meshes):
            raise BundleError(f"instance refers to missing geometry {index}")
        instances.append(InstanceData(
            mesh=index, transform=column_major_matrix(_member(instance, "transform")),
            visibility=visibility_mask(_member(instance, "visibility")),
            shadow_terminator_geometry_offset=max(
                _number(_member(instance, "shadow_terminator_geometry_offset")), 0.0),
            name=_text(_member(instance, "name"))))
    if any(0 in meshes[i.mesh].material_slots for i in instances):
        diagnostics.append(Diagnostic(
            "warning", "the Cycles default surface (Principled BSDF, base colour 0.8) is drawn "
                       "as a Diffuse BSDF of the same colour"))
    lights = []
    for light in _member(root, "lights") or []:
        tree = _member(light, "node_tree")
        if isinstance(tree, dict) and _member(tree, "surface_root") is not None:
            diagnostics.append(Diagnostic(
                "error", f"light '{_text(_member(light, 'name'))}' has a shader; light shaders "
                         "are not supported"))
        lights.append(_light(light))
    world_color = (0.0, 0.0, 0.0)
    world_visibility = abi.VIS_ALL
    world = _member(root, "world")
    if isinstance(world, dict):
        world_visibility = visibility_mask(_member(world, "visibility"))
        tree = _member(world, "node_tree")
        if isinstance(tree, dict) and _member(tree, "surface_root") is not None:
            program = lower_surface("__world__" + _text(_member(world, "name")), world, diagnostics)
            constant = _constant_emission(program)
            if constant is None:
                diagnostics.append(Diagnostic("error", "only a constant Background world is supported"))
            else:
                color, strength = constant
                world_color = tuple(c * strength for c in color)
        else:
            world_color = _float3(_member(world, "color"), (0.05, 0.05, 0.05))
    if _member(root, "world_environment") is not None:
        diagnostics.append(Diagnostic("error", "environment maps are not supported"))
    result.material_names = [m.name for m in materials]
    return SceneSnapshot(materials=materials, meshes=meshes, instances=instances,
                         camera=camera_data, world_color=world_color,
                         world_visibility=world_visibility, lights=lights,
                         name=_text(_member(root, "name"), os.path.basename(os.path.normpath(directory))))
