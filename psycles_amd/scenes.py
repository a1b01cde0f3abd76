"""Scene snapshots for the render path and the synthetic benchmark scenes.

``SceneSnapshot`` is the Python form of contract::SceneSnapshot
(include/psycles/contract/scene.h:458-489) restricted to what this render
path consumes: triangle meshes, instances, SurfaceProgram materials with
parameter blocks, background lights, one perspective camera and a constant
world colour.  ``to_desc`` lowers it to the flat ``psx_scene_desc`` of
include/psx.h.

The generators build the workloads the render path is measured on:

* ``cornell_box``: 640x400 Cornell box, 34 diffuse triangles plus a
  0.25 x 0.25 emissive quad (strength 15).
* ``instanced_icospheres``: a displaced geodesic icosphere of 20 f^2
  triangles (f = 224 gives 1,003,520), instanced 64 times with random
  rotation, scale and translation, each with a mix(diffuse, transparent)
  material, plus a ground quad and a constant world background.

All randomness comes from SplitMix64 streams with fixed seeds, so every
snapshot is bit-for-bit reproducible.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) as a scalar stream."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniforms(self, n: int) -> np.ndarray:
        """n doubles in [0, 1): the same values as n calls of uniform()."""
        idx = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            self.state = int((np.uint64(self.state) + np.uint64(n) *
                              np.uint64(0x9E3779B97F4A7C15)))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


# --------------------------------------------------------------------------
# SurfaceProgram construction
# --------------------------------------------------------------------------

@dataclass
class Material:
    """A SurfaceProgram (closure instructions + root) and its parameter block."""

    instructions: list  # tuples (op, a, b, factor, color, strength, roughness)
    root: int
    parameters: np.ndarray
    emission_sampling: int = abi.EMISSION_SAMPLING_AUTOMATIC
    name: str = ""

    def structure(self) -> tuple:
        return (tuple(self.instructions), self.root, len(self.parameters))


class ProgramBuilder:
    """Builds closure trees; children are always emitted before parents."""

    def __init__(self):
        self.instructions = []
        self.parameters: list[float] = []

    def _param(self, *values: float) -> int:
        index = len(self.parameters)
        self.parameters.extend(float(v) for v in values)
        return index

    def _emit(self, op, a=0, b=0, factor=0, color=0, strength=0, roughness=0) -> int:
        self.instructions.append((op, a, b, factor, color, strength, roughness))
        return len(self.instructions) - 1

    def diffuse(self, color, roughness=0.0) -> int:
        c = self._param(*color)
        r = self._param(roughness)
        return self._emit(abi.CLOSURE_DIFFUSE, color=c, roughness=r)

    def emission(self, color, strength) -> int:
        c = self._param(*color)
        s = self._param(strength)
        return self._emit(abi.CLOSURE_EMISSION, color=c, strength=s)

    def transparent(self, color=(1.0, 1.0, 1.0)) -> int:
        c = self._param(*color)
        return self._emit(abi.CLOSURE_TRANSPARENT, color=c)

    def add(self, a: int, b: int) -> int:
        return self._emit(abi.CLOSURE_ADD, a=a, b=b)

    def mix(self, a: int, b: int, factor: float) -> int:
        f = self._param(factor)
        return self._emit(abi.CLOSURE_MIX, a=a, b=b, factor=f)

    def build(self, root: int, emission_sampling=abi.EMISSION_SAMPLING_AUTOMATIC,
              name="") -> Material:
        return Material(list(self.instructions), root,
                        np.asarray(self.parameters, dtype=np.float32),
                        emission_sampling, name)


def diffuse_material(color, roughness=0.0, name="") -> Material:
    b = ProgramBuilder()
    return b.build(b.diffuse(color, roughness), name=name)


def emission_material(color, strength, name="") -> Material:
    b = ProgramBuilder()
    return b.build(b.emission(color, strength), name=name)


def mix_diffuse_transparent(albedo, factor, name="") -> Material:
    b = ProgramBuilder()
    root = b.mix(b.diffuse(albedo), b.transparent((1.0, 1.0, 1.0)), factor)
    return b.build(root, name=name)


def emission_diffuse_material(color, strength, albedo, name="") -> Material:
    b = ProgramBuilder()
    root = b.add(b.emission(color, strength), b.diffuse(albedo))
    return b.build(root, name=name)


# --------------------------------------------------------------------------
# snapshot
# --------------------------------------------------------------------------

def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32)


def translation(x, y, z) -> np.ndarray:
    m = identity()
    m[:3, 3] = (x, y, z)
    return m


def rotation_y(degrees: float) -> np.ndarray:
    a = math.radians(degrees)
    m = identity()
    m[0, 0] = math.cos(a)
    m[0, 2] = math.sin(a)
    m[2, 0] = -math.sin(a)
    m[2, 2] = math.cos(a)
    return m


def scale(x, y, z) -> np.ndarray:
    return np.diag([x, y, z, 1.0]).astype(np.float32)


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Camera-to-world transform; the camera looks down its local -z axis."""
    eye = np.asarray(eye, dtype=np.float64)
    forward = np.asarray(target, dtype=np.float64) - eye
    forward /= np.linalg.norm(forward)
    right = np.cross(forward, np.asarray(up, dtype=np.float64))
    right /= np.linalg.norm(right)
    true_up = np.cross(right, forward)
    m = np.eye(4)
    m[:3, 0] = right
    m[:3, 1] = true_up
    m[:3, 2] = -forward
    m[:3, 3] = eye
    return m.astype(np.float32)


def column_major(m: np.ndarray) -> list[float]:
    """Row-major numpy 4x4 -> Mat4f elements (column-major)."""
    return [float(v) for v in np.asarray(m, dtype=np.float32).T.reshape(-1)]


@dataclass
class MeshData:
    positions: np.ndarray            # (V, 3) float32
    triangles: np.ndarray            # (T, 3) uint32
    material_slots: list
    triangle_material_slots: np.ndarray | None = None
    normals: np.ndarray | None = None
    smooth: np.ndarray | None = None
    name: str = ""


@dataclass
class InstanceData:
    mesh: int
    transform: np.ndarray            # row-major 4x4
    material_overrides: list = field(default_factory=list)
    visibility: int = abi.VIS_ALL
    shadow_terminator_geometry_offset: float = 0.0
    name: str = ""


@dataclass
class CameraData:
    transform: np.ndarray
    field_of_view: float = math.radians(45.0)
    near_clip: float = 1.0e-4
    far_clip: float = 1.0e5
    sensor_fit: int = 1              # 0 horizontal, 1 vertical, 2 automatic
    projection: int = 0              # 0 perspective, 1 orthographic
    orthographic_scale: float = 1.0
    horizontal_field_of_view: float | None = None   # defaults to field_of_view


def material_structs(materials, keep: list):
    """psx_material array of `materials`; the buffers it points to go to `keep`."""
    mats = (abi.Material * max(len(materials), 1))()
    for i, m in enumerate(materials):
        instr = (abi.ClosureInstr * max(len(m.instructions), 1))()
        for j, t in enumerate(m.instructions):
            instr[j] = abi.ClosureInstr(*t, 0)
        keep.append(instr)
        mats[i].instructions = C.cast(instr, C.POINTER(abi.ClosureInstr))
        mats[i].instruction_count = len(m.instructions)
        mats[i].root = m.root
        params = np.ascontiguousarray(m.parameters.astype(np.float32))
        keep.append(params)
        mats[i].parameters = params.ctypes.data_as(C.POINTER(C.c_float))
        mats[i].parameter_count = len(m.parameters)
        mats[i].emission_sampling = m.emission_sampling
    keep.append(mats)
    return mats


@dataclass
class SceneSnapshot:
    materials: list
    meshes: list
    instances: list
    camera: CameraData
    world_color: tuple = (0.0, 0.0, 0.0)
    world_visibility: int = abi.VIS_ALL
    lights: list = field(default_factory=list)
    name: str = ""

    def triangle_count(self) -> int:
        return sum(int(self.meshes[i.mesh].triangles.shape[0]) for i in self.instances)

    def to_desc(self):
        """Lower to psx_scene_desc.  Returns (desc, keepalive)."""
        keep = []

        def arr(a, ctype):
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        mats = material_structs(self.materials, keep)
        meshes = (abi.Mesh * max(len(self.meshes), 1))()
        for i, m in enumerate(self.meshes):
            meshes[i].positions = arr(m.positions.astype(np.float32).reshape(-1), C.c_float)
            meshes[i].vertex_count = int(m.positions.shape[0])
            meshes[i].triangles = arr(m.triangles.astype(np.uint32).reshape(-1), C.c_uint32)
            meshes[i].triangle_count = int(m.triangles.shape[0])
            if m.normals is not None:
                meshes[i].normals = arr(m.normals.astype(np.float32).reshape(-1), C.c_float)
            if m.smooth is not None:
                meshes[i].triangle_smooth = arr(m.smooth.astype(np.uint8), C.c_uint8)
            meshes[i].material_slots = arr(np.asarray(m.material_slots, dtype=np.uint32), C.c_uint32)
            meshes[i].material_slot_count = len(m.material_slots)
            if m.triangle_material_slots is not None:
                meshes[i].triangle_material_slots = arr(
                    m.triangle_material_slots.astype(np.uint32), C.c_uint32)
        keep.append(meshes)
        insts = (abi.Instance * max(len(self.instances), 1))()
        for i, inst in enumerate(self.instances):
            insts[i].mesh = inst.mesh
            insts[i].visibility_mask = inst.visibility
            insts[i].transform[:] = column_major(inst.transform)
            if inst.material_overrides:
                insts[i].material_overrides = arr(
                    np.asarray(inst.material_overrides, dtype=np.uint32), C.c_uint32)
                insts[i].material_override_count = len(inst.material_overrides)
            insts[i].shadow_terminator_geometry_offset = inst.shadow_terminator_geometry_offset
        keep.append(insts)
        lights = (abi.Light * max(len(self.lights), 1))()
        for i, l in enumerate(self.lights):
            lights[i] = l
        keep.append(lights)
        d = abi.SceneDesc()
        d.abi_version = abi.PSX_ABI_VERSION
        d.materials = C.cast(mats, C.POINTER(abi.Material))
        d.material_count = len(self.materials)
        d.meshes = C.cast(meshes, C.POINTER(abi.Mesh))
        d.mesh_count = len(self.meshes)
        d.instances = C.cast(insts, C.POINTER(abi.Instance))
        d.instance_count = len(self.instances)
        d.lights = C.cast(lights, C.POINTER(abi.Light))
        d.light_count = len(self.lights)
        cam = self.camera
        d.camera.projection = cam.projection
        d.camera.sensor_fit = cam.sensor_fit
        d.camera.transform[:] = column_major(cam.transform)
        d.camera.field_of_view = cam.field_of_view
        d.camera.horizontal_field_of_view = (cam.field_of_view if cam.horizontal_field_of_view is None
                                             else cam.horizontal_field_of_view)
        d.camera.orthographic_scale = cam.orthographic_scale
        d.camera.near_clip = cam.near_clip
        d.camera.far_clip = cam.far_clip
        d.world_color[:] = [float(c) for c in self.world_color]
        d.world_visibility_mask = self.world_visibility
        return d, keep


# --------------------------------------------------------------------------
# geometry helpers
# --------------------------------------------------------------------------

def box_mesh(size, material_slot=0) -> tuple[np.ndarray, np.ndarray]:
    """Axis-aligned box centred at the origin: 8 vertices, 12 triangles with
    outward winding."""
    sx, sy, sz = (0.5 * s for s in size)
    p = np.array([[-sx, -sy, -sz], [sx, -sy, -sz], [sx, sy, -sz], [-sx, sy, -sz],
                  [-sx, -sy, sz], [sx, -sy, sz], [sx, sy, sz], [-sx, sy, sz]],
                 dtype=np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2],      # -z
                  [4, 5, 6], [4, 6, 7],      # +z
                  [0, 1, 5], [0, 5, 4],      # -y
                  [3, 7, 6], [3, 6, 2],      # +y
                  [0, 4, 7], [0, 7, 3],      # -x
                  [1, 2, 6], [1, 6, 5]],     # +x
                 dtype=np.uint32)
    return p, t


def quad(corners) -> tuple[np.ndarray, np.ndarray]:
    p = np.asarray(corners, dtype=np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    return p, t


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------

def cornell_box() -> SceneSnapshot:
    """Procedural Cornell box: 5 walls, a short and a tall box, 34 triangles,
    and a 0.25 x 0.25 emissive ceiling quad at y = 0.999 (strength 15)."""
    grey = diffuse_material((0.73, 0.73, 0.73), name="grey")
    red = diffuse_material((0.65, 0.05, 0.05), name="red")
    green = diffuse_material((0.12, 0.45, 0.15), name="green")
    light = emission_material((1.0, 1.0, 1.0), 15.0, name="light")
    materials = [grey, red, green, light]
    # walls of the cube [-1, 1]^3 without the +z face, winding facing inward
    wall_p = np.array([
        [-1, -1, -1], [1, -1, -1], [1, -1, 1], [-1, -1, 1],     # floor
        [-1, 1, -1], [-1, 1, 1], [1, 1, 1], [1, 1, -1],         # ceiling
        [-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1],     # back
        [-1, -1, -1], [-1, -1, 1], [-1, 1, 1], [-1, 1, -1],     # left
        [1, -1, -1], [1, 1, -1], [1, 1, 1], [1, -1, 1],         # right
    ], dtype=np.float32)
    wall_t = []
    for q in range(5):
        b = 4 * q
        wall_t += [[b, b + 2, b + 1], [b, b + 3, b + 2]]
    wall_t = np.asarray(wall_t, dtype=np.uint32)
    wall_slots = np.array([0, 0, 0, 0, 0, 0, 1, 1, 2, 2], dtype=np.uint32)
    walls = MeshData(wall_p, wall_t, [0, 1, 2], wall_slots, name="walls")
    sp, st = box_mesh((0.6, 0.6, 0.6))
    short_box = MeshData(sp, st, [0], name="short_box")
    tp, tt = box_mesh((0.6, 1.2, 0.6))
    tall_box = MeshData(tp, tt, [0], name="tall_box")
    h = 0.125
    lp, lt = quad([[-h, 0.999, -h], [h, 0.999, -h], [h, 0.999, h], [-h, 0.999, h]])
    light_quad = MeshData(lp, lt, [3], name="light")
    meshes = [walls, short_box, tall_box, light_quad]
    instances = [
        InstanceData(0, identity(), name="walls"),
        InstanceData(1, translation(0.33, -0.7, 0.38) @ rotation_y(-18.0), name="short_box"),
        InstanceData(2, translation(-0.33, -0.4, -0.3) @ rotation_y(15.0), name="tall_box"),
        InstanceData(3, identity(), name="light"),
    ]
    camera = CameraData(translation(0.0, 0.0, 3.9), math.radians(40.0), 1.0e-3, 100.0)
    return SceneSnapshot(materials, meshes, instances, camera, (0.0, 0.0, 0.0),
                         name="cornell")


def lamp(kind: int, transform, color=(1.0, 1.0, 1.0), power=1.0, size=0.0, spot_angle=math.pi / 4,
         spot_smooth=0.15, angle=0.0, normalize=True, sphere=True, use_mis=True,
         visibility=abi.VIS_ALL, max_bounces=1024) -> abi.Light:
    """contract::LightDesc for a point, spot, distant or background lamp
    (include/psycles/contract/scene.h:377-403); `transform` is 4x4 row-major
    numpy, its z axis the lamp's axis (spots shine along -z, the distant
    lamp's +z points at the sun)."""
    l = abi.Light()
    l.type = kind
    l.visibility_mask = visibility
    for i, v in enumerate(column_major(np.asarray(transform, dtype=np.float32))):
        l.transform[i] = v
    for i in range(3):
        l.color[i] = color[i]
    l.power = power
    l.size = size
    l.size_y = 0.0
    l.spread = math.pi
    l.spot_angle = spot_angle
    l.spot_smooth = spot_smooth
    l.angle = angle
    l.flags = ((abi.LIGHT_FLAG_NORMALIZE if normalize else 0) | (abi.LIGHT_FLAG_SPHERE if sphere else 0) |
               (abi.LIGHT_FLAG_USE_MIS if use_mis else 0) | abi.LIGHT_FLAG_CAST_SHADOW)
    l.max_bounces = max_bounces
    return l


def frame_z(direction) -> np.ndarray:
    """Rotation whose z axis is the unit `direction`; x and y complete a
    right-handed orthonormal frame."""
    z = np.asarray(direction, dtype=np.float64)
    z = z / np.linalg.norm(z)
    helper = np.array([0.0, 1.0, 0.0]) if abs(z[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    x = np.cross(helper, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    m = identity()
    m[:3, 0], m[:3, 1], m[:3, 2] = x, y, z
    return m


def cornell_lamps(with_quad: bool = True) -> SceneSnapshot:
    """The Cornell box lit by a sphere point lamp (r 0.08), a sphere spot
    lamp (r 0.05) aimed at the floor and a distant sun (angle 0.1 rad)
    entering through the open +z side; the ceiling quad is kept optionally."""
    snap = cornell_box()
    if not with_quad:
        snap.instances = snap.instances[:3]
    snap.lights = [
        lamp(abi.LIGHT_POINT, translation(0.45, 0.55, 0.2), (1.0, 0.9, 0.8), 12.0, size=0.08),
        lamp(abi.LIGHT_SPOT, translation(-0.5, 0.8, 0.4) @ frame_z((0.3, 1.0, -0.2)), (0.6, 0.8, 1.0), 25.0,
             size=0.05, spot_angle=0.9, spot_smooth=0.3),
        lamp(abi.LIGHT_DISTANT, frame_z((0.2, 0.5, 1.0)), (1.0, 1.0, 0.9), 2.0, angle=0.1),
    ]
    snap.name = "cornell_lamps"
    return snap


def cornell_settings(width=640, height=400) -> abi.RenderSettings:
    """16 spp, max 4 bounces, RR from bounce 3 (min_bounces 1), seed 0."""
    s = abi.default_settings(width, height)
    s.seed = 0
    i = s.integrator
    i.max_bounces = 4
    i.min_bounces = 1
    i.diffuse_bounces = 4
    i.glossy_bounces = 4
    i.transmission_bounces = 4
    return s


def geodesic_icosphere(frequency: int, displacement: float = 0.02,
                       seed: int = 2) -> tuple[np.ndarray, np.ndarray]:
    """Geodesic icosphere with 20 * frequency^2 triangles and 10 f^2 + 2 shared
    vertices, each displaced radially by U[-displacement, displacement]
    (SplitMix64, seed 2, in vertex order)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    ico_v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                      [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                      [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    ico_v /= np.linalg.norm(ico_v, axis=1, keepdims=True)
    ico_f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                      [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                      [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                      [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    f = frequency
    # barycentric grid points (i, j) with i + j <= f, local index per face
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    mask = ii + jj <= f
    gi = ii[mask]
    gj = jj[mask]
    local = -np.ones((f + 1, f + 1), dtype=np.int64)
    local[gi, gj] = np.arange(gi.size)
    a = ico_v[ico_f[:, 0]][:, None, :]
    b = ico_v[ico_f[:, 1]][:, None, :]
    c = ico_v[ico_f[:, 2]][:, None, :]
    u = (gi / f)[None, :, None]
    v = (gj / f)[None, :, None]
    pts = a + (b - a) * u + (c - a) * v          # (20, G, 3) on the flat faces
    pts = pts.reshape(-1, 3)
    key = np.round(pts * (1 << 20)).astype(np.int64)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inverse = inverse.reshape(-1)
    # keep first-occurrence order so the vertex numbering is deterministic
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(order.size)
    vert_index = remap[inverse]
    positions = pts[first[order]]
    positions /= np.linalg.norm(positions, axis=1, keepdims=True)
    rng = SplitMix64(seed)
    disp = (rng.uniforms(positions.shape[0]) * 2.0 - 1.0) * displacement
    positions *= (1.0 + disp)[:, None]
    # triangles per face: up (i,j),(i+1,j),(i,j+1) and down (i+1,j),(i+1,j+1),(i,j+1)
    up_i, up_j = np.nonzero((np.add.outer(np.arange(f), np.arange(f))) <= f - 1)
    dn_i, dn_j = np.nonzero((np.add.outer(np.arange(f), np.arange(f))) <= f - 2)
    tri_local = np.concatenate([
        np.stack([local[up_i, up_j], local[up_i + 1, up_j], local[up_i, up_j + 1]], 1),
        np.stack([local[dn_i + 1, dn_j], local[dn_i + 1, dn_j + 1], local[dn_i, dn_j + 1]], 1),
    ])
    g = gi.size
    tris = (tri_local[None, :, :] + (np.arange(20) * g)[:, None, None]).reshape(-1, 3)
    tris = vert_index[tris].astype(np.uint32)
    return positions.astype(np.float32), tris


def random_rotation(rng: SplitMix64) -> np.ndarray:
    """Uniform random rotation from a unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r
    return m


def instanced_icospheres(frequency: int = 224, instances: int = 64,
                         box: float = 40.0) -> SceneSnapshot:
    """Displaced icosphere instanced with random rotation, scale U[0.5, 2] and
    translation in a box^3 cube (SplitMix64 seed 2002); 64 one-structure
    mix(diffuse(albedo U[0,1]^3), transparent(white), fac U[0,1]) materials;
    a grey 0.5 ground quad at y = -22 and a constant 0.05 world background."""
    positions, tris = geodesic_icosphere(frequency)
    rng = SplitMix64(2002)
    materials = []
    inst = []
    for k in range(instances):
        albedo = (rng.uniform(), rng.uniform(), rng.uniform())
        fac = rng.uniform()
        materials.append(mix_diffuse_transparent(albedo, fac, name=f"mix{k}"))
    sphere = MeshData(positions, tris, [0], name="icosphere")
    half = box / 2.0
    for k in range(instances):
        rot = random_rotation(rng)
        s = 0.5 + 1.5 * rng.uniform()
        tx, ty, tz = ((rng.uniform() * 2.0 - 1.0) * half for _ in range(3))
        m = translation(tx, ty, tz).astype(np.float64) @ rot @ np.diag([s, s, s, 1.0])
        inst.append(InstanceData(0, m.astype(np.float32), material_overrides=[k],
                                 name=f"sphere{k}"))
    ground_mat = diffuse_material((0.5, 0.5, 0.5), name="ground")
    materials.append(ground_mat)
    g = 200.0
    gp, gt = quad([[-g, -22.0, -g], [-g, -22.0, g], [g, -22.0, g], [g, -22.0, -g]])
    ground = MeshData(gp, gt, [len(materials) - 1], name="ground")
    inst.append(InstanceData(1, identity(), name="ground"))
    camera = CameraData(look_at((0.0, 6.0, 58.0), (0.0, -2.0, 0.0)), math.radians(42.0),
                        1.0e-3, 1000.0)
    return SceneSnapshot(materials, [sphere, ground], inst, camera, (0.05, 0.05, 0.05),
                         name="instanced")


def instanced_settings(width=1920, height=1080) -> abi.RenderSettings:
    """64 spp, max 8 bounces, seed 0; all other integrator settings default."""
    s = abi.default_settings(width, height)
    s.seed = 0
    s.integrator.max_bounces = 8
    return s


def many_lights(frequency: int = 224, instances: int = 64, emissive: int = 10000, points: int = 256,
                spots: int = 64, box: float = 40.0) -> SceneSnapshot:
    """The instanced icospheres plus emissive triangles and analytic lamps.

    `emissive` random (instance, triangle) pairs (SplitMix64 seed 3) are
    copied to world space, offset 1e-4 along their geometric normal into one
    extra mesh and given add(emission, diffuse(0.6)) materials: strength is
    log-uniform in [1, 100], bucketed into 16 log-spaced materials of one
    structure.  `points` sphere point lamps, `spots` sphere spot lamps and one
    distant sun get random positions, directions, sizes and colours
    (SplitMix64 seed 33).  Meant to be rendered with the light tree."""
    snap = instanced_icospheres(frequency, instances, box)
    sphere = snap.meshes[0]
    spheres = [i for i in snap.instances if i.mesh == 0]
    rng = SplitMix64(3)
    buckets = 16
    first_material = len(snap.materials)
    for b in range(buckets):
        strength = 100.0 ** ((b + 0.5) / buckets)
        snap.materials.append(emission_diffuse_material((1.0, 0.85, 0.7), strength, (0.6, 0.6, 0.6),
                                                        name=f"glow{b}"))
    pos = np.empty((emissive, 3, 3), dtype=np.float64)
    slots = np.empty(emissive, dtype=np.uint32)
    tri_count = sphere.triangles.shape[0]
    for e in range(emissive):
        inst = spheres[min(int(rng.uniform() * len(spheres)), len(spheres) - 1)]
        tri = sphere.triangles[min(int(rng.uniform() * tri_count), tri_count - 1)]
        m = inst.transform.astype(np.float64)
        local = sphere.positions[tri].astype(np.float64)
        world = local @ m[:3, :3].T + m[:3, 3]
        n = np.cross(world[1] - world[0], world[2] - world[0])
        n /= np.linalg.norm(n)
        pos[e] = world + 1.0e-4 * n
        u = rng.uniform()
        slots[e] = min(int(u * buckets), buckets - 1)
    glow = MeshData(pos.reshape(-1, 3).astype(np.float32),
                    np.arange(3 * emissive, dtype=np.uint32).reshape(-1, 3),
                    list(range(first_material, first_material + buckets)),
                    triangle_material_slots=slots, name="glow")
    snap.meshes.append(glow)
    snap.instances.append(InstanceData(len(snap.meshes) - 1, identity(), name="glow"))
    rng = SplitMix64(33)
    half = box / 2.0

    def colour():
        return tuple(0.3 + 0.7 * rng.uniform() for _ in range(3))

    lights = []
    for _ in range(points):
        p = [(rng.uniform() * 2.0 - 1.0) * half * 1.2 for _ in range(3)]
        c = colour()
        power = 50.0 * 10.0 ** rng.uniform()
        size = 0.05 + 0.45 * rng.uniform()
        lights.append(lamp(abi.LIGHT_POINT, translation(*p), c, power, size=size))
    for _ in range(spots):
        p = [(rng.uniform() * 2.0 - 1.0) * half, 5.0 + 20.0 * rng.uniform(), (rng.uniform() * 2.0 - 1.0) * half]
        d = (rng.uniform() - 0.5, -1.0, rng.uniform() - 0.5)
        c = colour()
        power = 200.0 * 10.0 ** rng.uniform()
        size = 0.05 + 0.25 * rng.uniform()
        angle = 0.3 + 0.9 * rng.uniform()
        smooth = 0.4 * rng.uniform()
        # the spot shines along -z of its frame
        lights.append(lamp(abi.LIGHT_SPOT, translation(*p) @ frame_z((-d[0], -d[1], -d[2])), c, power,
                           size=size, spot_angle=angle, spot_smooth=smooth))
    lights.append(lamp(abi.LIGHT_DISTANT, frame_z((0.3, 1.0, 0.2)), (1.0, 0.95, 0.85), 1.5, angle=0.05))
    snap.lights = lights
    snap.name = "many_lights"
    return snap


def many_light_settings(width=1920, height=1080) -> abi.RenderSettings:
    """32 spp, max 6 bounces, seed 0, light-tree NEE."""
    s = abi.default_settings(width, height)
    s.seed = 0
    s.integrator.max_bounces = 6
    s.integrator.use_light_tree = 1
    return s


# --------------------------------------------------------------------------
# uneven-path batch: stacked transparent slabs over diffuse geometry
# --------------------------------------------------------------------------

@dataclass
class UnevenRequest:
    """One request of the uneven-path batch: its own scene, a square
    resolution and a sample count.  `volume` and `sigma_t` describe the
    optional homogeneous box the request would carry; this render path has no
    volume transport, so the box is drawn (the stream stays the same) but not
    built into the scene."""

    index: int
    snapshot: SceneSnapshot
    side: int
    spp: int
    slabs: int
    volume: bool
    sigma_t: float

    def settings(self, side: int | None = None) -> abi.RenderSettings:
        return uneven_settings(self.side if side is None else side, self.index)


def uneven_slabs(transparencies, albedos, boxes) -> SceneSnapshot:
    """A stack of 4 x 4 slabs at z = 1.5 ... -0.5 facing the camera, each a
    mix(diffuse(albedo), transparent(white), transparency) quad, in front of
    a grey 0.6 ground at y = -1 carrying three diffuse boxes; a 0.3 constant
    world and a distant sun (angle 0.05 rad, strength 3)."""
    materials = [diffuse_material((0.6, 0.6, 0.6), name="ground")]
    meshes = []
    inst = []
    g = 20.0
    gp, gt = quad([[-g, -1.0, -g], [-g, -1.0, g], [g, -1.0, g], [g, -1.0, -g]])
    meshes.append(MeshData(gp, gt, [0], name="ground"))
    inst.append(InstanceData(0, identity(), name="ground"))
    bp, bt = box_mesh((0.8, 0.8, 0.8))
    meshes.append(MeshData(bp, bt, [0], name="box"))
    for k, (albedo, (x, z, angle)) in enumerate(zip(albedos[:3], boxes)):
        materials.append(diffuse_material(albedo, name=f"box{k}"))
        inst.append(InstanceData(1, translation(x, -0.6, z) @ rotation_y(angle),
                                 material_overrides=[len(materials) - 1], name=f"box{k}"))
    n = len(transparencies)
    for k, t in enumerate(transparencies):
        z = 1.5 - 2.0 * k / max(n - 1, 1) if n > 1 else 0.5
        materials.append(mix_diffuse_transparent(albedos[3 + k], t, name=f"slab{k}"))
        sp, st = quad([[-2.0, -0.9, z], [2.0, -0.9, z], [2.0, 2.0, z], [-2.0, 2.0, z]])
        meshes.append(MeshData(sp, st, [len(materials) - 1], name=f"slab{k}"))
        inst.append(InstanceData(len(meshes) - 1, identity(), name=f"slab{k}"))
    camera = CameraData(look_at((0.0, 0.4, 4.2), (0.0, -0.2, -1.5)), math.radians(45.0), 1.0e-3, 100.0)
    snap = SceneSnapshot(materials, meshes, inst, camera, (0.3, 0.3, 0.3), name="uneven")
    snap.lights = [lamp(abi.LIGHT_DISTANT, frame_z((0.4, 1.0, 0.6)), (1.0, 0.95, 0.9), 3.0, angle=0.05)]
    return snap


def uneven_batch(count: int = 512, seed: int = 5) -> list:
    """The uneven-path batch: per request 1-16 slabs with transparency
    U[0.5, 0.95], slab and box albedos U[0.2, 0.9]^3, box placements, a
    homogeneous box with probability 0.5 (sigma_t U[0.1, 2]), a square
    resolution side U{128..1024} and U{1..256} samples (SplitMix64 `seed`)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        slabs = 1 + int(rng.uniform() * 16)
        transparencies = [0.5 + 0.45 * rng.uniform() for _ in range(slabs)]
        albedos = [tuple(0.2 + 0.7 * rng.uniform() for _ in range(3)) for _ in range(3 + slabs)]
        boxes = [(-1.5 + 3.0 * rng.uniform(), -2.5 + 1.5 * rng.uniform(), 90.0 * rng.uniform())
                 for _ in range(3)]
        volume = rng.uniform() < 0.5
        sigma_t = 0.1 + 1.9 * rng.uniform()
        side = 128 + int(rng.uniform() * 897)
        spp = 1 + int(rng.uniform() * 256)
        out.append(UnevenRequest(i, uneven_slabs(transparencies, albedos, boxes), side, spp, slabs,
                                 volume, sigma_t))
    return out


def uneven_settings(side: int, seed: int) -> abi.RenderSettings:
    """Square `side` x `side`, seed = request index, max 8 bounces with RR
    (min_bounces 0), 32 transparent bounces so every slab of a 16-slab stack
    is crossed by camera and shadow rays."""
    s = abi.default_settings(side, side)
    s.seed = seed
    s.integrator.max_bounces = 8
    s.integrator.transparent_max_bounces = 32
    return s
