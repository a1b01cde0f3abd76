"""Diagnostic path traces in the "psycles.cycles-path-trace" v2 layout.

`LuisaPathTrace` / `LuisaPathTraceRequest` counterparts
(include/psycles/luisa/path_tracer.h:20-43) over psx_session_path_trace:
one pixel (``pixel_y`` counted from the lower row of the full image, Cycles'
film convention) and one sample, 328 float4 slots laid out as
include/psycles/luisa/path_trace_schema.h:10-121 defines them.  The device
writes the global slots 0-6 and, for the first 4 surface events, every event
slot 0-47 the path reaches (state, hit and surface; light sample, NEE
evaluation, light shader and contribution,
src/luisa/path_kernel_direct_light_trace.cpp; closure pick and BSDF sample,
post-scatter state and MIS, src/luisa/path_kernel_surface_scatter.cpp), the
closure slots 48-71 (index, Cycles closure type and sample weight; weight;
normal) and the shadow-event slots 296-327 (shadow ray, range and
cast-shadow, source and light identities, first hit, transmittance); w = 1
marks a written slot.  Lamps have no object or shader index in this ABI:
their light_id object is 4294967295 and their light_shader the halves of
0xffffffff, what the reference writes for a light without them.  `write_raw_path_trace` writes the
JSON document of examples/render_blender_scene.cpp:77-112
("psycles.cycles-path-trace-raw"); `decode` names the written slots.
"""
from __future__ import annotations

import json
from dataclasses import dataclass

import numpy as np

from . import abi

SCHEMA_NAME = "psycles.cycles-path-trace"
RAW_SCHEMA_NAME = "psycles.cycles-path-trace-raw"
VERSION = abi.PATH_TRACE_VERSION
SLOT_COUNT = abi.PATH_TRACE_SLOTS
GLOBAL_SLOT_COUNT = 8
EVENT_SLOT_COUNT = 72
SHADOW_EVENT_SLOT_COUNT = 8
SHADOW_EVENT_BASE = 296
MAX_EVENTS = 4
MAX_CLOSURES = 8

GLOBAL_SLOTS = ("header", "rng", "filter", "lens_time", "ray_p", "ray_d", "ray_range", "reserved")
EVENT_SLOTS = (
    "state_depth", "state_lobes", "state_visibility", "state_flags", "throughput", "ray_p", "ray_d",
    "ray_range", "isect_coord", "isect_id", "surface_meta", "surface_p", "surface_ng", "surface_n",
    "random_scalars", "random_light", "random_bsdf", "light_meta", "light_id", "light_pdf", "light_d",
    "light_p", "light_ng", "light_eval", "closure_pick", "closure_random", "closure_weight", "closure_n",
    "bsdf_meta", "bsdf_wo", "bsdf_eval", "bsdf_roughness_eta", "post_depth", "post_throughput",
    "post_ray_p", "post_ray_d", "post_flags", "post_mis", "surface_flags", "post_visibility",
    "light_shader", "nee_bsdf", "nee_diffuse", "nee_glossy", "nee_weighted_bsdf", "nee_light_shader",
    "nee_unshadowed", "nee_contribution")
CLOSURE_BASE = 48
SHADOW_EVENT_SLOTS = ("shadow_ray_p", "shadow_ray_d", "shadow_ray_range", "shadow_source", "shadow_light",
                      "shadow_hit_id", "shadow_hit_coord", "shadow_transmittance")


def index(event: int, slot: str) -> int:
    """path_trace_schema::index(event, EventSlot)."""
    return GLOBAL_SLOT_COUNT + event * EVENT_SLOT_COUNT + EVENT_SLOTS.index(slot)


def shadow_index(event: int, slot: str) -> int:
    return SHADOW_EVENT_BASE + event * SHADOW_EVENT_SLOT_COUNT + SHADOW_EVENT_SLOTS.index(slot)


def closure_index(event: int, closure: int, field: int) -> int:
    return GLOBAL_SLOT_COUNT + event * EVENT_SLOT_COUNT + CLOSURE_BASE + closure * 3 + field


def uint32_from_halves(lo: float, hi: float) -> int:
    """Inverse of trace_uint32 (src/luisa/path_kernel_setup.cpp:430-433)."""
    return int(lo) | (int(hi) << 16)


@dataclass
class PathTrace:
    pixel_x: int
    pixel_y: int
    sample: int
    slots: np.ndarray          # (328, 4) float32


@dataclass
class PathTraceRequest:
    pixel_x: int
    pixel_y: int
    sample: int
    sink: object = None        # anything with write(PathTrace)


def trace_path(session, request: PathTraceRequest, total: int) -> PathTrace:
    """Trace `request`'s pixel and sample of a render of `total` samples on
    `session` and hand the record to its sink."""
    slots, _ = session.path_trace(request.pixel_x, request.pixel_y, request.sample, total)
    trace = PathTrace(request.pixel_x, request.pixel_y, request.sample, slots)
    if request.sink is not None:
        request.sink.write(trace)
    return trace


def write_raw_path_trace(trace: PathTrace, path) -> None:
    document = {"schema": RAW_SCHEMA_NAME, "trace_schema_version": VERSION, "pixel_x": trace.pixel_x,
                "pixel_y": trace.pixel_y, "sample": trace.sample,
                "slots": [[float(v) for v in slot] for slot in trace.slots]}
    with open(path, "w") as f:
        json.dump(document, f)


def read_raw_path_trace(path) -> PathTrace:
    with open(path) as f:
        document = json.load(f)
    if document.get("schema") != RAW_SCHEMA_NAME or document.get("trace_schema_version") != VERSION:
        raise ValueError("not a version 2 raw path trace")
    slots = np.asarray(document["slots"], dtype=np.float32)
    if slots.shape != (SLOT_COUNT, 4):
        raise ValueError("a raw path trace holds 328 slots of 4 values")
    return PathTrace(int(document["pixel_x"]), int(document["pixel_y"]), int(document["sample"]), slots)


def decode(trace: PathTrace) -> dict:
    """Written slots by name: {"global": {...}, "events": [{...}, ...]}."""
    s = trace.slots
    out = {"global": {name: tuple(float(v) for v in s[i, :3])
                      for i, name in enumerate(GLOBAL_SLOTS) if s[i, 3] == 1.0}, "events": []}
    for e in range(MAX_EVENTS):
        fields = {}
        for j, name in enumerate(EVENT_SLOTS):
            i = GLOBAL_SLOT_COUNT + e * EVENT_SLOT_COUNT + j
            if s[i, 3] == 1.0:
                fields[name] = tuple(float(v) for v in s[i, :3])
        if not fields:
            break
        if "state_flags" in fields:
            lo, hi, _ = fields["state_flags"]
            fields["path_flag"] = uint32_from_halves(lo, hi)
        if "state_visibility" in fields:
            _, lo, hi = fields["state_visibility"]
            fields["path_visibility"] = uint32_from_halves(lo, hi)
        out["events"].append(fields)
    return out
