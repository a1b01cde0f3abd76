"""ctypes mirror of include/psx.h, the C ABI of the render path.

The structs below match the header field for field; ``load_library`` opens the
in-tree ``libpsx.so`` and fails loudly when it has not been built, so no caller
can silently fall back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PSX_ABI_VERSION = 2

PSX_OK = 0
PSX_ERROR_INVALID_ARGUMENT = -1
PSX_ERROR_DEVICE = -2
PSX_ERROR_OUT_OF_MEMORY = -3
PSX_ERROR_CANCELLED = -4
PSX_ERROR_RANGE = -5
PSX_ERROR_INCOMPLETE = -6
PSX_ERROR_UNSUPPORTED = -7
PSX_ERROR_NO_DEVICE = -8

STATUS_NAMES = {
    0: "ok", -1: "invalid argument", -2: "device error", -3: "out of memory",
    -4: "cancelled", -5: "bad sample range", -6: "incomplete samples",
    -7: "unsupported", -8: "no device",
}

# contract::PassKind values (include/psycles/contract/render.h:15-42)
PASS_COMBINED = 0
PASS_DEPTH = 1
PASS_NORMAL = 2
PASS_POSITION = 3
PASS_ALBEDO = 4
PASS_EMISSION = 5
PASS_DIFFUSE_DIRECT = 6
PASS_DIFFUSE_INDIRECT = 7
PASS_GLOSSY_DIRECT = 8
PASS_GLOSSY_INDIRECT = 9
PASS_TRANSMISSION_DIRECT = 10
PASS_TRANSMISSION_INDIRECT = 11
PASS_VOLUME_DIRECT = 12
PASS_VOLUME_INDIRECT = 13
PASS_GLOSSY_COLOR = 23
PASS_TRANSMISSION_COLOR = 24
#: the passes accumulated from light-pass sums (LightPassBuffer)
LIGHT_PASS_KINDS = (PASS_DIFFUSE_DIRECT, PASS_DIFFUSE_INDIRECT, PASS_GLOSSY_DIRECT, PASS_GLOSSY_INDIRECT,
                    PASS_TRANSMISSION_DIRECT, PASS_TRANSMISSION_INDIRECT, PASS_VOLUME_DIRECT, PASS_VOLUME_INDIRECT,
                    PASS_EMISSION, 25, PASS_GLOSSY_COLOR, PASS_TRANSMISSION_COLOR)
PASS_DENOISING_NORMAL = 18
PASS_DENOISING_ALBEDO = 19
PASS_SAMPLE_COUNT = 20
PASS_ENVIRONMENT = 25

FILTER_BOX = 0
FILTER_GAUSSIAN = 1
FILTER_BLACKMAN_HARRIS = 2

DIRECT_LIGHT_MIS = 0
DIRECT_LIGHT_FORWARD = 1
DIRECT_LIGHT_NEE = 2

EMISSION_SAMPLING_NONE = 0
EMISSION_SAMPLING_AUTOMATIC = 1
EMISSION_SAMPLING_FRONT = 2
EMISSION_SAMPLING_BACK = 3
EMISSION_SAMPLING_FRONT_BACK = 4

VIS_CAMERA = 1 << 0
VIS_DIFFUSE = 1 << 1
VIS_GLOSSY = 1 << 2
VIS_TRANSMISSION = 1 << 3
VIS_SHADOW = 1 << 4
VIS_VOLUME_SCATTER = 1 << 5
VIS_ALL = 0x3F

CLOSURE_NULL = 0
CLOSURE_ADD = 1
CLOSURE_MIX = 2
CLOSURE_DIFFUSE = 3
CLOSURE_EMISSION = 4
CLOSURE_TRANSPARENT = 5

LIGHT_POINT = 0
LIGHT_SPOT = 1
LIGHT_AREA = 2
LIGHT_DISTANT = 3
LIGHT_BACKGROUND = 4

LIGHT_FLAG_NORMALIZE = 1
LIGHT_FLAG_ELLIPSE = 2
LIGHT_FLAG_SPHERE = 4
LIGHT_FLAG_USE_MIS = 8
LIGHT_FLAG_CAST_SHADOW = 16

INTEGRATOR_MEGAKERNEL = 0
INTEGRATOR_WAVEFRONT = 1
BACKEND_COUNT_TRAVERSAL = 1
BACKEND_NO_GRAPHS = 2
BACKEND_TIME_TRACE = 4
BACKEND_GENERATED_PROGRAMS = 8
BACKEND_BVH_WIDTH_2 = 16
BACKEND_BVH_WIDTH_3 = 32
PROGRAM_LIGHT_PASSES = 1
PROGRAM_LIGHTS = 2

TRACE_EVENTS = 8
TRACE_WORDS = 8
PATH_TRACE_SLOTS = 328
PATH_TRACE_VERSION = 2
TRACE_FLAGS = 0
TRACE_RNG_OFFSET = 1
TRACE_OBJECT = 2
TRACE_PRIMITIVE = 3
TRACE_LIGHT = 4
TRACE_CLOSURE = 5
TRACE_TERMINATE_BITS = 6
TRACE_RR_BITS = 7

u32 = C.c_uint32
f32 = C.c_float
P = C.POINTER


class ClosureInstr(C.Structure):
    _fields_ = [("op", u32), ("a", u32), ("b", u32), ("factor", u32),
                ("color", u32), ("strength", u32), ("roughness", u32),
                ("reserved", u32)]


class Material(C.Structure):
    _fields_ = [("instructions", P(ClosureInstr)), ("instruction_count", u32),
                ("root", u32), ("parameters", P(f32)),
                ("parameter_count", u32), ("emission_sampling", u32)]


class Mesh(C.Structure):
    _fields_ = [("positions", P(f32)), ("vertex_count", u32),
                ("triangles", P(u32)), ("triangle_count", u32),
                ("normals", P(f32)), ("triangle_smooth", P(C.c_uint8)),
                ("material_slots", P(u32)), ("material_slot_count", u32),
                ("triangle_material_slots", P(u32))]


class Instance(C.Structure):
    _fields_ = [("mesh", u32), ("visibility_mask", u32),
                ("transform", f32 * 16), ("material_overrides", P(u32)),
                ("material_override_count", u32),
                ("shadow_terminator_geometry_offset", f32)]


class Light(C.Structure):
    _fields_ = [("type", u32), ("visibility_mask", u32),
                ("transform", f32 * 16), ("color", f32 * 3), ("power", f32),
                ("size", f32), ("size_y", f32), ("spread", f32),
                ("spot_angle", f32), ("spot_smooth", f32), ("angle", f32),
                ("flags", u32), ("max_bounces", u32)]


class Camera(C.Structure):
    _fields_ = [("projection", u32), ("sensor_fit", u32),
                ("transform", f32 * 16), ("field_of_view", f32),
                ("horizontal_field_of_view", f32), ("lens_shift_x", f32),
                ("lens_shift_y", f32), ("near_clip", f32), ("far_clip", f32),
                ("orthographic_scale", f32)]


class SceneDesc(C.Structure):
    _fields_ = [("abi_version", u32), ("materials", P(Material)),
                ("material_count", u32), ("meshes", P(Mesh)),
                ("mesh_count", u32), ("instances", P(Instance)),
                ("instance_count", u32), ("lights", P(Light)),
                ("light_count", u32), ("camera", Camera),
                ("world_color", f32 * 3), ("world_visibility_mask", u32)]


class IntegratorSettings(C.Structure):
    _fields_ = [("max_bounces", u32), ("min_bounces", u32),
                ("diffuse_bounces", u32), ("glossy_bounces", u32),
                ("transmission_bounces", u32), ("volume_bounces", u32),
                ("transparent_min_bounces", u32),
                ("transparent_max_bounces", u32),
                ("sample_clamp_direct", f32), ("sample_clamp_indirect", f32),
                ("filter_glossy", f32), ("film_exposure", f32),
                ("light_sampling_threshold", f32),
                ("reflective_caustics", u32), ("refractive_caustics", u32),
                ("use_light_tree", u32), ("direct_light_sampling", u32)]


class RenderSettings(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("window_x", u32),
                ("window_y", u32), ("window_width", u32),
                ("window_height", u32), ("seed", u32),
                ("transparent_background", u32), ("pixel_filter", u32),
                ("filter_width", f32), ("pass_alpha_threshold", f32),
                ("integrator", IntegratorSettings), ("passes", P(u32)),
                ("pass_channels", P(u32)), ("pass_count", u32)]


SINK_BEGIN = C.CFUNCTYPE(None, C.c_void_p, P(RenderSettings))
SINK_WRITE = C.CFUNCTYPE(None, C.c_void_p, u32, u32, u32, u32, u32, u32,
                         u32, u32, P(f32))
SINK_END = C.CFUNCTYPE(None, C.c_void_p, C.c_int)
DIAG_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)


class Sink(C.Structure):
    _fields_ = [("user", C.c_void_p), ("begin", SINK_BEGIN),
                ("write", SINK_WRITE), ("end", SINK_END)]


class BackendOptions(C.Structure):
    _fields_ = [("next_event_estimation", u32),
                ("max_samples_per_dispatch", u32), ("integrator", u32),
                ("flags", u32)]


class RenderStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("samples", C.c_uint64),
                ("dispatches", C.c_uint64), ("kernel_ms", C.c_double),
                ("node_visits", C.c_uint64), ("triangle_tests", C.c_uint64),
                ("trace_ms", C.c_double), ("trace_launches", C.c_uint64),
                ("graph_launches", C.c_uint64), ("graph_builds", C.c_uint64)]


class WavefrontTuning(C.Structure):
    _fields_ = [("tile_width", u32), ("refill_lanes", u32), ("trace_chunk", u32),
                ("queue_chunk", u32), ("list_chunk", u32), ("sort_shading", u32), ("xcd_segments", u32)]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class TuningReport(C.Structure):
    _fields_ = [("chosen", WavefrontTuning), ("default_ms", f32), ("chosen_ms", f32),
                ("candidates", u32), ("probe_samples", u32)]


class ServerOptions(C.Structure):
    _fields_ = [("path_capacity", u32), ("streams", u32), ("hbm_budget", C.c_uint64),
                ("viewport_max_pixels", u32), ("viewport_max_samples", u32), ("target_batch_ms", u32),
                ("reserved", u32 * 3)]


REQUEST_QUEUED, REQUEST_RUNNING, REQUEST_DONE, REQUEST_CANCELLED, REQUEST_FAILED = range(5)


class RequestStatus(C.Structure):
    _fields_ = [("state", u32), ("samples_done", u32), ("batches", u32), ("viewport", u32),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("first_launch_ms", C.c_double),
                ("done_ms", C.c_double)]


class ServingStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("paths", C.c_uint64), ("rays", C.c_uint64), ("edits", C.c_uint64),
                ("hbm_in_use", C.c_uint64), ("peak_hbm", C.c_uint64), ("busy_ms", C.c_double)]


INVALID_ID = 0xFFFFFFFF
INTERSECT_NOALLOC_MAX_QUERIES = 1 << 30
QUERY_CLOSEST = 0
QUERY_SHADOW = 1


class RayQuery(C.Structure):
    _fields_ = [("origin", f32 * 3), ("tmin", f32), ("direction", f32 * 3), ("tmax", f32),
                ("visibility", u32), ("source_object", u32), ("source_primitive", u32),
                ("light_object", u32), ("light_primitive", u32), ("kind", u32),
                ("transparent_depth", u32), ("transparent_max", u32)]


class RayHit(C.Structure):
    _fields_ = [("t", f32), ("u", f32), ("v", f32), ("object", u32), ("primitive", u32), ("hit", u32),
                ("traversals", u32), ("reserved", u32), ("transmittance", f32 * 3), ("reserved2", u32)]


assert C.sizeof(RayQuery) == 64 and C.sizeof(RayHit) == 48

# psx_probe_function values and record widths (floats in, floats out)
PROBE_FILM_CLAMP = 0
PROBE_T_OFFSET = 1
PROBE_SELF_EXCLUSION = 2
PROBE_SHADOW_TERMINATOR = 3
PROBE_TRIANGLE_SAMPLE = 4
PROBE_TRIANGLE_PDF = 5
PROBE_DIFFUSE_EVAL = 6
PROBE_SURFACE_SHADOW_ORIGIN = 7
PROBE_AREA_SAMPLE = 8
PROBE_VOLUME = 9
PROBE_WIDTHS = {PROBE_FILM_CLAMP: (6, 3), PROBE_T_OFFSET: (1, 1), PROBE_SELF_EXCLUSION: (21, 4),
                PROBE_SHADOW_TERMINATOR: (47, 4), PROBE_TRIANGLE_SAMPLE: (14, 10), PROBE_TRIANGLE_PDF: (15, 3),
                PROBE_DIFFUSE_EVAL: (13, 4), PROBE_SURFACE_SHADOW_ORIGIN: (59, 4), PROBE_AREA_SAMPLE: (21, 10),
                PROBE_VOLUME: (32, 24)}


def make_queries(records) -> "C.Array":
    """RayQuery array from dicts with keys origin, direction, tmin, tmax and
    optional visibility, source, light, kind, transparent_depth,
    transparent_max."""
    arr = (RayQuery * len(records))()
    for q, r in zip(arr, records):
        q.origin[:] = [float(x) for x in r["origin"]]
        q.direction[:] = [float(x) for x in r["direction"]]
        q.tmin = float(r.get("tmin", 0.0))
        q.tmax = float(r.get("tmax", 3.4028234663852886e38))
        q.visibility = int(r.get("visibility", VIS_ALL))
        q.source_object, q.source_primitive = r.get("source", (INVALID_ID, INVALID_ID))
        q.light_object, q.light_primitive = r.get("light", (INVALID_ID, INVALID_ID))
        q.kind = int(r.get("kind", QUERY_CLOSEST))
        q.transparent_depth = int(r.get("transparent_depth", 0))
        q.transparent_max = int(r.get("transparent_max", 8))
    return arr


def default_settings(width: int, height: int) -> RenderSettings:
    """contract::RenderSettings defaults (include/psycles/contract/render.h:81-117)."""
    s = RenderSettings()
    s.width = width
    s.height = height
    s.window_width = width
    s.window_height = height
    s.pixel_filter = FILTER_BOX
    s.filter_width = 1.0
    s.pass_alpha_threshold = 0.5
    i = s.integrator
    i.max_bounces = 12
    i.min_bounces = 0
    i.diffuse_bounces = 4
    i.glossy_bounces = 4
    i.transmission_bounces = 12
    i.volume_bounces = 0
    i.transparent_min_bounces = 0
    i.transparent_max_bounces = 8
    i.film_exposure = 1.0
    i.light_sampling_threshold = 0.01
    i.reflective_caustics = 1
    i.refractive_caustics = 1
    i.use_light_tree = 0
    i.direct_light_sampling = DIRECT_LIGHT_MIS
    return s


PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "libpsx.so"

#: every symbol include/psx.h declares
EXPORTED_SYMBOLS = [
    "psx_version", "psx_device_count", "psx_last_error", "psx_render_settings_default",
    "psx_backend_create", "psx_backend_destroy", "psx_scene_compile",
    "psx_scene_destroy", "psx_scene_update_parameters",
    "psx_session_create", "psx_session_destroy", "psx_session_reset",
    "psx_session_render", "psx_session_cancel", "psx_session_stats",
    "psx_session_render_device", "psx_session_read_film",
    "psx_session_primary_hits", "psx_session_trace_events", "psx_session_path_trace",
    "psx_sobol_table", "psx_session_sampler_probe", "psx_session_read_light_passes",
    "psx_server_create", "psx_server_destroy", "psx_server_submit",
    "psx_server_update_parameters", "psx_server_status", "psx_server_wait",
    "psx_server_read_film", "psx_server_write_passes", "psx_server_cancel",
    "psx_server_release", "psx_server_stats",
    "psx_scene_intersect", "psx_scene_intersect_device", "psx_scene_intersect_device_noalloc",
    "psx_probe_input_width", "psx_probe_output_width", "psx_backend_probe",
    "psx_wavefront_tuning_default", "psx_session_set_tuning", "psx_session_get_tuning",
    "psx_session_autotune", "psx_surface_program_source", "psx_surface_program_build",
    "psx_backend_generated_programs",
]

_lib = None


def load_library(path: os.PathLike | None = None):
    """Open libpsx.so (built by __graft_entry__.build()) and declare prototypes."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    lib_path = Path(path) if path is not None else LIBRARY_PATH
    if not lib_path.exists():
        raise RuntimeError(
            f"{lib_path} is missing: build the HIP render library first "
            "(python -c 'import __graft_entry__; __graft_entry__.build()')")
    lib = C.CDLL(str(lib_path))
    vp = C.c_void_p
    lib.psx_version.restype = C.c_char_p
    lib.psx_device_count.restype = C.c_int
    lib.psx_last_error.restype = C.c_char_p
    lib.psx_render_settings_default.argtypes = [P(RenderSettings)]
    lib.psx_backend_create.argtypes = [C.c_int, P(BackendOptions), P(vp)]
    lib.psx_backend_destroy.argtypes = [vp]
    lib.psx_scene_compile.argtypes = [vp, P(SceneDesc), DIAG_FN, vp, P(vp)]
    lib.psx_scene_destroy.argtypes = [vp]
    lib.psx_scene_update_parameters.argtypes = [vp, u32, P(f32), u32]
    lib.psx_session_create.argtypes = [vp, P(RenderSettings), P(vp)]
    lib.psx_session_destroy.argtypes = [vp]
    lib.psx_session_reset.argtypes = [vp, P(RenderSettings)]
    lib.psx_session_render.argtypes = [vp, u32, u32, u32, u32, P(Sink)]
    lib.psx_session_cancel.argtypes = [vp]
    lib.psx_session_stats.argtypes = [vp, P(RenderStats)]
    lib.psx_session_render_device.argtypes = [vp, u32, u32, u32]
    lib.psx_session_read_film.argtypes = [vp, P(f32), P(f32), P(f32), P(u32)]
    lib.psx_session_primary_hits.argtypes = [vp, u32, u32, P(u32)]
    lib.psx_session_trace_events.argtypes = [vp, u32, u32, P(u32)]
    lib.psx_session_path_trace.argtypes = [vp, u32, u32, u32, u32, P(f32), P(u32)]
    lib.psx_sobol_table.argtypes = [u32, P(f32), P(u32)]
    lib.psx_session_sampler_probe.argtypes = [vp, u32, u32, u32, u32, P(u32)]
    lib.psx_session_read_light_passes.argtypes = [vp, P(f32)]
    lib.psx_session_read_light_passes.restype = C.c_int
    lib.psx_server_create.argtypes = [vp, P(ServerOptions), P(vp)]
    lib.psx_server_destroy.argtypes = [vp]
    lib.psx_server_submit.argtypes = [vp, vp, P(RenderSettings), u32, u32, u32, P(C.c_uint64)]
    lib.psx_server_update_parameters.argtypes = [vp, vp, u32, P(f32), u32]
    lib.psx_server_status.argtypes = [vp, C.c_uint64, P(RequestStatus)]
    lib.psx_server_wait.argtypes = [vp, C.c_uint64, C.c_double]
    lib.psx_server_read_film.argtypes = [vp, C.c_uint64, P(f32), P(f32), P(f32), P(u32)]
    lib.psx_server_write_passes.argtypes = [vp, C.c_uint64, P(Sink)]
    lib.psx_server_cancel.argtypes = [vp, C.c_uint64]
    lib.psx_server_release.argtypes = [vp, C.c_uint64]
    lib.psx_server_stats.argtypes = [vp, P(ServingStats)]
    lib.psx_scene_intersect.argtypes = [vp, P(RayQuery), u32, P(RayHit)]
    lib.psx_scene_intersect.restype = C.c_int
    lib.psx_scene_intersect_device.argtypes = [vp, vp, u32, vp]
    lib.psx_scene_intersect_device.restype = C.c_int
    lib.psx_scene_intersect_device_noalloc.argtypes = [vp, vp, u32, vp]
    lib.psx_scene_intersect_device_noalloc.restype = C.c_int
    lib.psx_probe_input_width.argtypes = [u32]
    lib.psx_probe_input_width.restype = u32
    lib.psx_probe_output_width.argtypes = [u32]
    lib.psx_probe_output_width.restype = u32
    lib.psx_backend_probe.argtypes = [vp, u32, P(f32), u32, P(f32)]
    lib.psx_backend_probe.restype = C.c_int
    lib.psx_wavefront_tuning_default.argtypes = [P(WavefrontTuning)]
    lib.psx_wavefront_tuning_default.restype = None
    lib.psx_session_set_tuning.argtypes = [vp, P(WavefrontTuning)]
    lib.psx_session_set_tuning.restype = C.c_int
    lib.psx_session_get_tuning.argtypes = [vp, P(WavefrontTuning)]
    lib.psx_session_get_tuning.restype = C.c_int
    lib.psx_session_autotune.argtypes = [vp, u32, P(TuningReport)]
    lib.psx_session_autotune.restype = C.c_int
    lib.psx_surface_program_source.argtypes = [P(Material), u32, u32, C.c_char_p, C.c_size_t,
                                               P(C.c_size_t), P(u32)]
    lib.psx_surface_program_source.restype = C.c_int
    lib.psx_surface_program_build.argtypes = [P(Material), u32, u32, C.c_char_p, C.c_char_p, C.c_size_t,
                                              P(C.c_uint64)]
    lib.psx_surface_program_build.restype = C.c_int
    lib.psx_backend_generated_programs.argtypes = [vp]
    lib.psx_backend_generated_programs.restype = u32
    for name in ("psx_server_create", "psx_server_submit", "psx_server_update_parameters",
                 "psx_server_status", "psx_server_wait", "psx_server_read_film",
                 "psx_server_write_passes", "psx_server_cancel", "psx_server_release",
                 "psx_server_stats"):
        getattr(lib, name).restype = C.c_int
    lib.psx_server_destroy.restype = None
    for name in ("psx_backend_create", "psx_scene_compile",
                 "psx_scene_update_parameters", "psx_session_create",
                 "psx_session_reset", "psx_session_render",
                 "psx_session_stats", "psx_session_render_device",
                 "psx_session_read_film", "psx_session_primary_hits",
                 "psx_session_trace_events", "psx_session_path_trace", "psx_sobol_table",
                 "psx_session_sampler_probe"):
        getattr(lib, name).restype = C.c_int
    for name in ("psx_backend_destroy", "psx_scene_destroy",
                 "psx_session_destroy", "psx_session_cancel",
                 "psx_render_settings_default"):
        getattr(lib, name).restype = None
    if path is None:
        _lib = lib
    return lib


class PsxError(RuntimeError):
    def __init__(self, where: str, status: int):
        detail = ""
        if status in (PSX_ERROR_DEVICE, PSX_ERROR_OUT_OF_MEMORY) and _lib is not None:
            detail = (_lib.psx_last_error() or b"").decode(errors="replace")
        super().__init__(f"{where} failed: {STATUS_NAMES.get(status, status)} ({status})"
                         + (f": {detail}" if detail else ""))
        self.status = status


def check(where: str, status: int) -> None:
    if status != PSX_OK:
        raise PsxError(where, status)
