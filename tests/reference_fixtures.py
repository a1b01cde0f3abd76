"""Scenes, queries and probe records built from the reference's function
fixtures in golden/reference_vectors.json (scene_traversal, surface_ray,
film_light, triangle_light), shared by the oracle tests and the GPU tests.

The traversal fixture names Cycles object and primitive ids; the scene built
here keeps the reference's instance order (without its curve instance), so
the instance index order equals the object id order of every tied pair and
the tie rule (greatest object, then primitive) compares the same way."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from psycles_amd import abi, scenes

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_vectors.json"
FLT_MAX = float(np.finfo(np.float32).max)


def vectors() -> dict:
    with open(GOLDEN) as f:
        return json.load(f)


def row_major(elements) -> np.ndarray:
    """Mat4f elements (column-major) -> row-major 4x4."""
    return np.asarray(elements, dtype=np.float32).reshape(4, 4).T.copy()


def _diffuse_material():
    return scenes.diffuse_material((0.5, 0.5, 0.5))


class TraversalScene:
    """The fixture's triangle instances as a snapshot.  Meshes with one user
    are stored in world space by the scene compiler (the Cycles static
    transform rule), so an instance the fixture marks `applied` gets a mesh of
    its own, and one it marks not applied shares its mesh; an instance whose
    mesh would otherwise have one user shares it with an invisible instance
    appended after the fixture's instances."""

    def __init__(self):
        v = vectors()["scene_traversal"]
        self.spec = v
        meshes, mesh_of, users = [], {}, {}
        insts = []
        # meshes: one per applied instance, one per shared geometry
        for inst in v["instances"]:
            key = inst["geometry"]
            if key not in mesh_of:
                m = v["meshes"][inst["mesh"]]
                mesh_of[key] = len(meshes)
                meshes.append(scenes.MeshData(positions=np.asarray(m["vertices"], np.float32),
                                              triangles=np.asarray(m["triangles"], np.uint32),
                                              material_slots=[0]))
            users[key] = users.get(key, 0) + 1
        self.objects = []
        for inst in v["instances"]:
            insts.append(scenes.InstanceData(mesh=mesh_of[inst["geometry"]],
                                             transform=row_major(v["transforms"][inst["transform"]])))
            self.objects.append(inst)
        for inst in v["instances"]:
            key = inst["geometry"]
            if not inst["applied"] and users[key] == 1:
                insts.append(scenes.InstanceData(mesh=mesh_of[key], transform=scenes.identity(), visibility=0))
                users[key] += 1
        prim_base = np.cumsum([0] + [m.triangles.shape[0] for m in meshes])
        self.mesh_base = [int(prim_base[mesh_of[i["geometry"]]]) for i in v["instances"]]
        self.snapshot = scenes.SceneSnapshot(
            materials=[_diffuse_material()], meshes=meshes, instances=insts,
            camera=scenes.CameraData(transform=scenes.look_at((0, 0, 10), (0, 0, 0))))
        self.index_of_object = {inst["object"]: i for i, inst in enumerate(v["instances"])}

    def identity(self, cycles_object, cycles_primitive):
        """(instance, scene primitive) of a Cycles identity; invalid stays invalid."""
        if cycles_object not in self.index_of_object:
            return abi.INVALID_ID, abi.INVALID_ID
        i = self.index_of_object[cycles_object]
        if cycles_primitive == abi.INVALID_ID:
            return i, abi.INVALID_ID
        local = cycles_primitive - self.objects[i]["primitive_offset"]
        return i, self.mesh_base[i] + local

    def cycles(self, instance, primitive):
        """Cycles (object, primitive) of an instance index and scene primitive."""
        inst = self.objects[instance]
        return inst["object"], inst["primitive_offset"] + (primitive - self.mesh_base[instance])

    def queries(self):
        out = []
        for q in self.spec["queries"]:
            r = {"origin": q["origin"], "direction": q["direction"], "tmin": 0.0, "tmax": q["tmax"],
                 "visibility": 0xFF}
            if "source" in q:
                r["source"] = self.identity(*q["source"])
            if "light" in q:
                r["light"] = self.identity(*q["light"])
            out.append(r)
        return out


def _fma(a, b, c) -> float:
    """float32 fused multiply-add (the double product of two floats is exact)."""
    return float(np.float32(np.float64(np.float32(a)) * np.float64(np.float32(b)) + np.float64(np.float32(c))))


def check_traversal(scene: TraversalScene, hits, intersect) -> list[str]:
    """Failures of `hits` (psx_ray_hit per fixture query) against the expected
    records; `intersect(records)` runs follow-up queries (record 20's shadow
    ray from the reconstructed hit point)."""
    spec = scene.spec
    tol = spec["tolerance"]
    errors = []
    for q, h in zip(spec["queries"], hits):
        rec = q["record"]
        if q["expected"] is None:
            if h.hit:
                errors.append(f"record {rec}: expected a miss, got object {h.object}")
            continue
        t, obj, prim = q["expected"]
        if not h.hit:
            errors.append(f"record {rec}: missed")
            continue
        got = scene.cycles(h.object, h.primitive)
        if got != (obj, prim) or abs(h.t - t) > tol:
            errors.append(f"record {rec}: got t={h.t!r} {got}, expected t={t} ({obj}, {prim})")
        if "shadow" in q and h.hit:
            # The fixture rebuilds the hit point from the Cycles barycentrics
            # and casts a shadow ray that must not hit the near-coincident
            # sibling.  It takes the point to world space with the backend's
            # float4x4 product, whose rounding is the Luisa lowering; this
            # path reconstructs surface points with the Cycles FMA affine
            # (fma(x, c0, fma(y, c1, fma(z, c2, t)))), used here.  The
            # unfused column sum lands on the wrong side of the sibling
            # (a hit at t ~ 5e-10), so the record does tell them apart.
            inst = scene.objects[h.object]
            m = scene.spec["meshes"][inst["mesh"]]
            tri = m["triangles"][h.primitive - scene.mesh_base[h.object]]
            p = [np.asarray(m["vertices"][k], np.float32) for k in tri]
            bu, bv = np.float32(h.u), np.float32(h.v)
            obj_p = (p[0] + bu * (p[1] - p[0])) + bv * (p[2] - p[0])
            e = np.asarray(scene.spec["transforms"][inst["transform"]], np.float32).reshape(4, 4)
            world = [_fma(obj_p[0], e[0, k], _fma(obj_p[1], e[1, k], _fma(obj_p[2], e[2, k], e[3, k])))
                     for k in range(3)]
            s = q["shadow"]
            sh = intersect([{"origin": world, "direction": s["direction"], "tmin": 0.0,
                             "tmax": s["tmax"], "visibility": 0xFF,
                             "source": (h.object, h.primitive)}])[0]
            if s["expected_miss"] and sh.hit:
                errors.append(f"record {rec}: shadow ray hit object {scene.cycles(sh.object, sh.primitive)}")
    return errors


def ordered_hits_snapshot():
    """The surface-ray fixture's two-triangle mesh (the farther one first)."""
    v = vectors()["surface_ray"]["ordered_shadow_hits"]
    mesh = scenes.MeshData(positions=np.asarray(v["vertices"], np.float32),
                           triangles=np.asarray(v["triangles"], np.uint32), material_slots=[0])
    # a second user keeps the mesh in object space under the identity
    insts = [scenes.InstanceData(mesh=0, transform=scenes.identity()),
             scenes.InstanceData(mesh=0, transform=scenes.identity(), visibility=0)]
    return scenes.SceneSnapshot(materials=[_diffuse_material()], meshes=[mesh], instances=insts,
                                camera=scenes.CameraData(transform=scenes.look_at((0, 0, 10), (0, 0, 0))))


def check_ordered_hits(intersect) -> list[str]:
    """Closest-to-farthest iteration with the one-ulp tmin advance, and the
    explicit exclusion of (instance 0, primitive 1)."""
    v = vectors()["surface_ray"]["ordered_shadow_hits"]
    ray = {"origin": v["ray"]["origin"], "direction": v["ray"]["direction"], "tmin": 0.0,
           "tmax": v["ray"]["tmax"], "visibility": 0xFF}
    e = v["expected"]
    errors = []
    first = intersect([ray])[0]
    if not first.hit or first.t != e["first"][0] or first.primitive != e["first"][1]:
        errors.append(f"first hit {first.t} prim {first.primitive}")
    offset = float(np.frombuffer(np.array([first.t], np.float32).tobytes(), np.uint32)[0] + 1)
    tmin = float(np.array([int(offset)], np.uint32).view(np.float32)[0])
    second = intersect([dict(ray, tmin=tmin)])[0]
    if not second.hit or second.t != e["second"][0] or second.primitive != e["second"][1]:
        errors.append(f"second hit {second.t} prim {second.primitive}")
    tmin2 = float(np.array([np.array([second.t], np.float32).view(np.uint32)[0] + 1], np.uint32).view(np.float32)[0])
    third = intersect([dict(ray, tmin=tmin2)])[0]
    if third.hit:
        errors.append("third query hit")
    ex = intersect([dict(ray, source=tuple(e["excluded_source"]))])[0]
    if not ex.hit or ex.t != e["excluded_first"][0] or ex.primitive != e["excluded_first"][1] or ex.object != 0:
        errors.append(f"excluded query {ex.t} prim {ex.primitive}")
    return errors


def _special(x):
    return {"nan": float("nan"), "inf": float("inf"), "-inf": float("-inf"),
            "max": float(np.finfo(np.float32).max)}.get(x, x) if isinstance(x, str) else x


def film_clamp_inputs():
    recs = vectors()["film_light"]["records"]
    rows = [[*map(_special, r["contribution"]), r["depth"], r["direct"], r["indirect"]] for r in recs]
    return np.asarray(rows, np.float32), np.asarray([r["expected"] for r in recs], np.float32)


def self_exclusion_inputs():
    v = vectors()["surface_ray"]["self_exclusion"]
    rows = [[*v["world_origin"], *v["world_normal"], *c["object_origin"], *c["object_direction"],
             *sum(v["p"], [])] for c in v["cases"]]
    return np.asarray(rows, np.float32), [c["certified"] for c in v["cases"]]


IDENTITY_AFFINE = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]


def _normalize(v):
    v = np.asarray(v, np.float32)
    return v * (np.float32(1.0) / np.sqrt(np.float32(np.dot(v, v))))


def terminator_inputs(with_world_to_object: bool = False):
    """SHADOW_TERMINATOR (or SURFACE_SHADOW_ORIGIN) records of the fixture's
    four terminator cases."""
    v = vectors()["surface_ray"]["shadow_terminator"]
    n = np.asarray(v["n"], np.float32)
    ns = _normalize((n[0] + n[1] + n[2]) / np.float32(3.0))
    rows = []
    for c in v["cases"]:
        light = _normalize(c["light"])
        row = [*v["position"], *ns, *v["geometric_normal"], *light, c["offset"], float(c["smooth"]), 0.0,
               *v["barycentric"], *sum(v["p"], []), *sum(v["n"], []), *IDENTITY_AFFINE]
        if with_world_to_object:
            row += IDENTITY_AFFINE
        rows.append(row)
    return np.asarray(rows, np.float32), v


def neighbouring_inputs():
    """SURFACE_SHADOW_ORIGIN records of the fixture's flat-triangle cases."""
    v = vectors()["surface_ray"]
    p = np.asarray(v["shadow_terminator"]["p"], np.float32)
    light = _normalize([0.3, -0.2, 1.0])
    flat = [0.0, 0.0, 1.0]
    rows = []
    for c in v["neighbouring_triangles"]["cases"]:
        bu, bv = (np.float32(x) for x in c["barycentric"])
        pos = p[0] * (np.float32(1.0) - bu - bv) + p[1] * bu + p[2] * bv
        rows.append([*pos, *flat, *flat, *light, 0.0, 0.0, 0.0, bu, bv, *p.reshape(-1), *flat * 3,
                     *IDENTITY_AFFINE, *IDENTITY_AFFINE])
    return np.asarray(rows, np.float32)


def check_terminator(out, v) -> list[str]:
    errors = []
    pos = np.asarray(v["position"], np.float32)
    for i, c in enumerate(v["cases"]):
        if bool(out[i, 3]) != c["skip_self"]:
            errors.append(f"case {i}: skip_self {out[i, 3]}")
        if c["expected"] == "unchanged":
            if not np.array_equal(out[i, :3].view(np.uint32), pos.view(np.uint32)):
                errors.append(f"case {i}: origin moved to {out[i, :3]}")
        elif np.max(np.abs(out[i, :3] - np.asarray(c["expected"], np.float32))) > v["tolerance"]:
            errors.append(f"case {i}: {out[i, :3]} vs {c['expected']}")
    return errors


def check_neighbouring(out) -> list[str]:
    errors = []
    interior, edge = out[0], out[1]
    if abs(interior[0]) > 1e-6 or abs(interior[1] + 1.0 / 3.0) > 1e-6 or \
            np.float32(interior[2]).view(np.uint32) != 0 or interior[3] != 1.0:
        errors.append(f"interior origin {interior}")
    if abs(edge[0]) > 1e-6 or abs(edge[1] + 1.0) > 1e-6 or not edge[2] > 0.0 or edge[3] != 1.0:
        errors.append(f"shared-edge origin {edge}")
    return errors


def triangle_inputs():
    v = vectors()["triangle_light"]
    p = np.asarray(v["near"]["p"], np.float32)
    far = p.copy()
    far[:, 2] = v["far_z"]
    ref = v["near"]["reference"]
    rnd = v["near"]["random"]
    return np.asarray([[*ref, *p.reshape(-1), *rnd], [*ref, *far.reshape(-1), *rnd]], np.float32), p, far


def check_triangle(sample, pdf_at_sample, far_pdf_expected) -> list[str]:
    """sample: TRIANGLE_SAMPLE outputs (near, far); pdf_at_sample: TRIANGLE_PDF
    of the near sample's position."""
    v = vectors()["triangle_light"]
    tol = v["tolerance_relative"]

    def close(a, b):
        return abs(a - b) <= tol * max(1.0, abs(a), abs(b))

    errors = []
    near, far = sample
    if not (near[8] == 1.0 and near[9] == 1.0 and near[7] > 0.0):
        errors.append(f"near sample: valid {near[8]}, solid angle {near[9]}, pdf {near[7]}")
    if not (pdf_at_sample[1] == 1.0 and pdf_at_sample[2] == 1.0 and close(pdf_at_sample[0], near[7])):
        errors.append(f"forward pdf {pdf_at_sample} vs sampled {near[7]}")
    b = v["expected"]["far_collision"]["barycentric"]
    pos = np.array([-2 * (1 - b[0] - b[1]) + 2 * b[0], -2 * (1 - b[0] - b[1]) - 2 * b[0] + 2 * b[1], v["far_z"]])
    if not (far[9] == 0.0 and far[8] == 1.0 and close(far[7], far_pdf_expected)
            and all(close(float(far[i]), float(pos[i])) for i in range(3))):
        errors.append(f"far sample {far} (expected position {pos}, pdf {far_pdf_expected})")
    return errors


def far_pdf() -> float:
    v = vectors()["triangle_light"]
    d2 = 0.3 * 0.3 + 1.0 + v["far_z"] ** 2
    return float(d2 * np.sqrt(d2) / (8.0 * v["far_z"]))
