"""Python mirror of Psycles' render contract over the psx C ABI.

Same names and meanings as include/psycles/contract/render.h:119-181:

* ``RendererBackend.compile_scene(snapshot) -> SceneCompilation`` with
  ``scene`` (None on failure) and ``diagnostics``;
* ``RendererBackend.create_session(scene, settings) -> RenderSession`` or
  None for invalid settings (src/luisa/path_tracer.cpp:22-75);
* ``RenderSession.render_samples(SampleRange, OutputSink) -> bool`` - False
  on cancel, a bad range or a changed total, exactly like
  src/luisa/path_tracer_session.cpp:272-413;
* ``OutputSink.begin / write(PassTile) / end(cancelled)``.

Every call goes through libpsx.so; there is no other implementation.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class SampleRange:
    first: int = 0
    count: int = 0
    offset: int = 0
    total: int = 0


@dataclass
class PassTile:
    pass_kind: int
    channels: int
    window: tuple
    full_extent: tuple
    pixels: np.ndarray


class OutputSink:
    def begin(self, settings) -> None:
        pass

    def write(self, tile: PassTile) -> None:
        pass

    def end(self, cancelled: bool) -> None:
        pass


class MemoryOutputSink(OutputSink):
    """Keeps every pass as an (H, W, C) array (include/psycles/io/image.h)."""

    def __init__(self):
        self.passes: dict[int, np.ndarray] = {}
        self.began = False
        self.ended = None

    def begin(self, settings) -> None:
        self.began = True

    def write(self, tile: PassTile) -> None:
        self.passes[tile.pass_kind] = tile.pixels

    def end(self, cancelled: bool) -> None:
        self.ended = bool(cancelled)


@dataclass
class RenderDiagnostic:
    message: str


@dataclass
class SceneCompilation:
    scene: "CompiledScene | None"
    diagnostics: list = field(default_factory=list)

    def ok(self) -> bool:
        return self.scene is not None


class CompiledScene:
    def __init__(self, backend: "RendererBackend", handle, snapshot):
        self._backend = backend
        self.handle = handle
        self.snapshot = snapshot

    def update_parameters(self, material: int, parameters) -> bool:
        """Parameter-only rebind; False when the block size changed."""
        p = np.ascontiguousarray(parameters, dtype=np.float32)
        rc = self._backend.lib.psx_scene_update_parameters(
            self.handle, material, p.ctypes.data_as(C.POINTER(C.c_float)), p.size)
        return rc == abi.PSX_OK

    def intersect(self, records) -> list:
        """psx_ray_hit records for query dicts (abi.make_queries): closest
        hits and transparent-shadow transmittance through the device
        traversal."""
        queries = abi.make_queries(records)
        hits = (abi.RayHit * len(records))()
        abi.check("psx_scene_intersect",
                  self._backend.lib.psx_scene_intersect(self.handle, queries, len(records), hits))
        return list(hits)

    def close(self) -> None:
        if self.handle:
            self._backend.lib.psx_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SinkCallbacks:
    """The ctypes callbacks of one psx_sink.  ctypes prints and drops an
    exception raised inside a callback, so each callback stores the first one
    here and `reraise()` raises it once the C call has returned."""

    def __init__(self, output: OutputSink, settings):
        self.error: BaseException | None = None

        def guard(fn):
            def call(*args):
                if self.error is not None:
                    return
                try:
                    fn(*args)
                except BaseException as e:  # re-raised by reraise()
                    self.error = e
            return call

        def on_begin(user, settings_ptr):
            output.begin(settings)

        def on_write(user, kind, channels, wx, wy, ww, wh, fw, fh, pixels):
            arr = np.ctypeslib.as_array(pixels, shape=(wh * ww * channels,)).copy()
            output.write(PassTile(kind, channels, (wx, wy, ww, wh), (fw, fh), arr.reshape(wh, ww, channels)))

        def on_end(user, cancelled):
            output.end(bool(cancelled))

        self.refs = (abi.SINK_BEGIN(guard(on_begin)), abi.SINK_WRITE(guard(on_write)),
                     abi.SINK_END(guard(on_end)))

    def reraise(self) -> None:
        if self.error is not None:
            err, self.error = self.error, None
            raise err


def make_sink(output: OutputSink, settings):
    """psx_sink forwarding to `output`; keep the returned callbacks alive for
    the duration of the call that uses the sink and call their `reraise()`
    after it, so an exception raised by the sink reaches the caller."""
    callbacks = SinkCallbacks(output, settings)
    sink = abi.Sink()
    sink.begin, sink.write, sink.end = callbacks.refs
    return sink, callbacks


class RenderSession:
    def __init__(self, backend: "RendererBackend", scene: CompiledScene, handle, settings):
        self._backend = backend
        self.scene = scene
        self.handle = handle
        self.settings = settings
        self.last_status = abi.PSX_OK

    @property
    def window(self) -> tuple:
        s = self.settings
        return (s.window_width or s.width, s.window_height or s.height)

    def reset(self, settings) -> None:
        abi.check("psx_session_reset",
                  self._backend.lib.psx_session_reset(self.handle, C.byref(settings)))
        self.settings = settings

    def render_samples(self, samples: SampleRange, output: OutputSink) -> bool:
        lib = self._backend.lib
        sink, callbacks = make_sink(output, self.settings)
        rc = lib.psx_session_render(self.handle, samples.first, samples.count, samples.offset,
                                    samples.total, C.byref(sink))
        self.last_status = rc
        callbacks.reraise()
        if rc in (abi.PSX_ERROR_DEVICE, abi.PSX_ERROR_OUT_OF_MEMORY, abi.PSX_ERROR_NO_DEVICE):
            raise abi.PsxError("psx_session_render", rc)
        return rc == abi.PSX_OK

    def render_device(self, first: int, count: int, total: int) -> None:
        """Accumulate samples on the device without reading the film back."""
        abi.check("psx_session_render_device",
                  self._backend.lib.psx_session_render_device(self.handle, first, count, total))

    def cancel(self) -> None:
        self._backend.lib.psx_session_cancel(self.handle)

    @property
    def tuning(self) -> abi.WavefrontTuning:
        """The wavefront launch shape (tile width, refill lanes, chunks)."""
        t = abi.WavefrontTuning()
        abi.check("psx_session_get_tuning", self._backend.lib.psx_session_get_tuning(self.handle, C.byref(t)))
        return t

    def set_tuning(self, tuning: abi.WavefrontTuning | dict) -> None:
        if isinstance(tuning, dict):
            base = self.tuning
            for name, value in tuning.items():
                setattr(base, name, value)
            tuning = base
        abi.check("psx_session_set_tuning", self._backend.lib.psx_session_set_tuning(self.handle, C.byref(tuning)))

    def autotune(self, probe_samples: int = 4) -> abi.TuningReport:
        """Times probe dispatches of candidate launch shapes and keeps the
        fastest (the film does not depend on the shape)."""
        rep = abi.TuningReport()
        abi.check("psx_session_autotune",
                  self._backend.lib.psx_session_autotune(self.handle, probe_samples, C.byref(rep)))
        return rep

    def stats(self) -> abi.RenderStats:
        st = abi.RenderStats()
        abi.check("psx_session_stats", self._backend.lib.psx_session_stats(self.handle, C.byref(st)))
        return st

    def read_film(self):
        w, h = self.window
        combined = np.zeros((h, w, 4), dtype=np.float32)
        normal = np.zeros((h, w, 4), dtype=np.float32)
        albedo = np.zeros((h, w, 4), dtype=np.float32)
        counts = np.zeros((h, w), dtype=np.uint32)
        f = C.POINTER(C.c_float)
        abi.check("psx_session_read_film", self._backend.lib.psx_session_read_film(
            self.handle, combined.ctypes.data_as(f), normal.ctypes.data_as(f),
            albedo.ctypes.data_as(f), counts.ctypes.data_as(C.POINTER(C.c_uint32))))
        return combined, normal, albedo, counts

    def read_light_passes(self) -> np.ndarray:
        """Raw light-pass sums, (H, W, 12, 4) in LightPassBuffer order."""
        w, h = self.window
        out = np.zeros((h, w, 12, 4), dtype=np.float32)
        abi.check("psx_session_read_light_passes", self._backend.lib.psx_session_read_light_passes(
            self.handle, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def primary_hits(self, sample_index: int, total: int) -> np.ndarray:
        w, h = self.window
        out = np.zeros((h, w, 4), dtype=np.uint32)
        abi.check("psx_session_primary_hits", self._backend.lib.psx_session_primary_hits(
            self.handle, sample_index, total, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def trace_events(self, sample_index: int, total: int) -> np.ndarray:
        w, h = self.window
        out = np.zeros((h, w, abi.TRACE_EVENTS, abi.TRACE_WORDS), dtype=np.uint32)
        abi.check("psx_session_trace_events", self._backend.lib.psx_session_trace_events(
            self.handle, sample_index, total, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def path_trace(self, pixel_x: int, pixel_y: int, sample_index: int, total: int):
        """Schema v2 path trace of one pixel (`pixel_y` counted from the lower
        row of the full image) and sample: ((328, 4) float32 slots, (8, 8)
        uint32 discrete events)."""
        slots = np.zeros((abi.PATH_TRACE_SLOTS, 4), dtype=np.float32)
        records = np.zeros((abi.TRACE_EVENTS, abi.TRACE_WORDS), dtype=np.uint32)
        abi.check("psx_session_path_trace", self._backend.lib.psx_session_path_trace(
            self.handle, pixel_x, pixel_y, sample_index, total, slots.ctypes.data_as(C.POINTER(C.c_float)),
            records.ctypes.data_as(C.POINTER(C.c_uint32))))
        return slots, records

    def sampler_probe(self, sample_index: int, total: int, first_dim: int, count: int) -> np.ndarray:
        w, h = self.window
        out = np.zeros((h, w, 1 + 4 * count), dtype=np.uint32)
        abi.check("psx_session_sampler_probe", self._backend.lib.psx_session_sampler_probe(
            self.handle, sample_index, total, first_dim, count,
            out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def close(self) -> None:
        if self.handle:
            self._backend.lib.psx_session_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RendererBackend:
    """luisa_backend::LuisaPathTracerBackend counterpart on one MI355X."""

    def __init__(self, device: int = 0, next_event_estimation: bool = True,
                 max_samples_per_dispatch: int = 4, integrator: int = abi.INTEGRATOR_MEGAKERNEL,
                 count_traversal: bool = False, graphs: bool = True, time_trace: bool = False,
                 generated_programs: bool = False, bvh_width: int = 4):
        self.lib = abi.load_library()
        if bvh_width not in (2, 3, 4):
            raise ValueError("bvh_width is 2, 3 or 4")
        self.bvh_width = bvh_width
        flags = abi.BACKEND_COUNT_TRAVERSAL if count_traversal else 0
        if bvh_width == 2:
            flags |= abi.BACKEND_BVH_WIDTH_2
        elif bvh_width == 3:
            flags |= abi.BACKEND_BVH_WIDTH_3
        if generated_programs:
            flags |= abi.BACKEND_GENERATED_PROGRAMS
        if not graphs:
            flags |= abi.BACKEND_NO_GRAPHS
        if time_trace:
            flags |= abi.BACKEND_TIME_TRACE
        opts = abi.BackendOptions(1 if next_event_estimation else 0, max_samples_per_dispatch,
                                  integrator, flags)
        handle = C.c_void_p()
        abi.check("psx_backend_create", self.lib.psx_backend_create(device, C.byref(opts),
                                                                    C.byref(handle)))
        self.handle = handle

    @property
    def generated_programs(self) -> int:
        """Generated shade modules this backend has compiled."""
        return int(self.lib.psx_backend_generated_programs(self.handle))

    def compile_scene(self, snapshot) -> SceneCompilation:
        desc, keep = snapshot.to_desc()
        diagnostics: list[RenderDiagnostic] = []

        def on_diag(user, msg):
            diagnostics.append(RenderDiagnostic(msg.decode()))

        cb = abi.DIAG_FN(on_diag)
        handle = C.c_void_p()
        rc = self.lib.psx_scene_compile(self.handle, C.byref(desc), cb, None, C.byref(handle))
        del keep
        if rc != abi.PSX_OK:
            if rc in (abi.PSX_ERROR_DEVICE, abi.PSX_ERROR_OUT_OF_MEMORY):
                raise abi.PsxError("psx_scene_compile", rc)
            return SceneCompilation(None, diagnostics)
        return SceneCompilation(CompiledScene(self, handle, snapshot), diagnostics)

    def create_session(self, scene: CompiledScene, settings) -> RenderSession | None:
        handle = C.c_void_p()
        rc = self.lib.psx_session_create(scene.handle, C.byref(settings), C.byref(handle))
        if rc == abi.PSX_ERROR_INVALID_ARGUMENT:
            return None
        abi.check("psx_session_create", rc)
        return RenderSession(self, scene, handle, settings)

    def probe(self, function: int, inputs) -> np.ndarray:
        """psx_backend_probe over rows of `inputs` (widths in abi.PROBE_WIDTHS)."""
        wi, wo = abi.PROBE_WIDTHS[function]
        x = np.ascontiguousarray(np.atleast_2d(np.asarray(inputs, dtype=np.float32)))
        if x.shape[1] != wi or self.lib.psx_probe_input_width(function) != wi:
            raise ValueError(f"probe {function} takes {wi} floats per record")
        out = np.zeros((x.shape[0], wo), dtype=np.float32)
        abi.check("psx_backend_probe", self.lib.psx_backend_probe(
            self.handle, function, x.ctypes.data_as(C.POINTER(C.c_float)), x.shape[0],
            out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def close(self) -> None:
        if self.handle:
            self.lib.psx_backend_destroy(self.handle)
            self.handle = None


def with_passes(settings, passes=((abi.PASS_COMBINED, 4), (abi.PASS_NORMAL, 3),
                                  (abi.PASS_ALBEDO, 3), (abi.PASS_SAMPLE_COUNT, 1))):
    """Attach pass requests to RenderSettings; returns (settings, keepalive)."""
    kinds = np.asarray([p[0] for p in passes], dtype=np.uint32)
    chans = np.asarray([p[1] for p in passes], dtype=np.uint32)
    settings.passes = kinds.ctypes.data_as(C.POINTER(C.c_uint32))
    settings.pass_channels = chans.ctypes.data_as(C.POINTER(C.c_uint32))
    settings.pass_count = len(passes)
    return settings, (kinds, chans)
