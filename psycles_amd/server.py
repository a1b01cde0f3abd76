"""Render server: a request loop over resident compiled scenes.

The serving form of `RenderSession.render_samples`: requests (a SampleRange of
one RenderSettings on one CompiledScene) are queued, admitted within an HBM
budget, and rendered as chunks inside shared wavefront batches on several
streams, so requests that arrive while others run are batched with them.
Parameter-only material edits are ordered with submissions: a request sees
exactly the edits submitted before it.  Each request's film is bit-identical
to a session rendering the same range alone.  See `psx_server_*` in
include/psx.h.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from . import abi
from .render import CompiledScene, OutputSink, RendererBackend, make_sink


# servers still open at interpreter exit are closed before module teardown,
# while the scenes they reference are alive and no worker thread outlives them
_OPEN_SERVERS: "weakref.WeakSet[RenderServer]" = weakref.WeakSet()


@atexit.register
def _close_open_servers() -> None:
    for server in list(_OPEN_SERVERS):
        server.close()


class ServerRequest:
    def __init__(self, server: "RenderServer", request_id: int, settings):
        self.server = server
        self.id = request_id
        self.settings = settings

    @property
    def window(self) -> tuple:
        s = self.settings
        return (s.window_width or s.width, s.window_height or s.height)

    def status(self) -> abi.RequestStatus:
        st = abi.RequestStatus()
        abi.check("psx_server_status", self.server.lib.psx_server_status(self.server.handle, self.id, C.byref(st)))
        return st

    def wait(self, timeout_ms: float = -1.0) -> bool:
        """True when done; False on timeout or cancellation."""
        rc = self.server.lib.psx_server_wait(self.server.handle, self.id, float(timeout_ms))
        if rc in (abi.PSX_OK,):
            return True
        if rc in (abi.PSX_ERROR_INCOMPLETE, abi.PSX_ERROR_CANCELLED):
            return False
        raise abi.PsxError("psx_server_wait", rc)

    def latency_ms(self) -> float:
        st = self.status()
        return st.done_ms - st.submit_ms

    def read_film(self):
        w, h = self.window
        combined = np.zeros((h, w, 4), dtype=np.float32)
        normal = np.zeros((h, w, 4), dtype=np.float32)
        albedo = np.zeros((h, w, 4), dtype=np.float32)
        counts = np.zeros((h, w), dtype=np.uint32)
        f = C.POINTER(C.c_float)
        abi.check("psx_server_read_film", self.server.lib.psx_server_read_film(
            self.server.handle, self.id, combined.ctypes.data_as(f), normal.ctypes.data_as(f),
            albedo.ctypes.data_as(f), counts.ctypes.data_as(C.POINTER(C.c_uint32))))
        return combined, normal, albedo, counts

    def write_passes(self, output: OutputSink) -> None:
        sink, callbacks = make_sink(output, self.settings)
        rc = self.server.lib.psx_server_write_passes(self.server.handle, self.id, C.byref(sink))
        callbacks.reraise()
        abi.check("psx_server_write_passes", rc)

    def cancel(self) -> None:
        abi.check("psx_server_cancel", self.server.lib.psx_server_cancel(self.server.handle, self.id))

    def release(self) -> None:
        abi.check("psx_server_release", self.server.lib.psx_server_release(self.server.handle, self.id))


class RenderServer:
    def __init__(self, backend: RendererBackend, path_capacity: int = 0, streams: int = 0,
                 hbm_budget: int = 0, viewport_max_pixels: int = 0, viewport_max_samples: int = 0,
                 target_batch_ms: int = 0):
        self.backend = backend
        self.lib = backend.lib
        opts = abi.ServerOptions()
        opts.path_capacity = path_capacity
        opts.streams = streams
        opts.hbm_budget = hbm_budget
        opts.viewport_max_pixels = viewport_max_pixels
        opts.viewport_max_samples = viewport_max_samples
        opts.target_batch_ms = target_batch_ms
        handle = C.c_void_p()
        abi.check("psx_server_create", self.lib.psx_server_create(backend.handle, C.byref(opts), C.byref(handle)))
        self.handle = handle
        # the scenes of submitted requests and edits: a scene must outlive
        # every request on it, so the server holds them until it closes
        self._scenes = {}
        _OPEN_SERVERS.add(self)

    def submit(self, scene: CompiledScene, settings, first: int, count: int, total: int) -> ServerRequest:
        rid = C.c_uint64()
        abi.check("psx_server_submit", self.lib.psx_server_submit(
            self.handle, scene.handle, C.byref(settings), first, count, total, C.byref(rid)))
        self._scenes[id(scene)] = scene
        return ServerRequest(self, rid.value, settings)

    def update_parameters(self, scene: CompiledScene, material: int, parameters) -> None:
        p = np.ascontiguousarray(parameters, dtype=np.float32)
        abi.check("psx_server_update_parameters", self.lib.psx_server_update_parameters(
            self.handle, scene.handle, material, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))
        self._scenes[id(scene)] = scene

    def stats(self) -> abi.ServingStats:
        st = abi.ServingStats()
        abi.check("psx_server_stats", self.lib.psx_server_stats(self.handle, C.byref(st)))
        return st

    def close(self) -> None:
        if self.handle:
            self.lib.psx_server_destroy(self.handle)
            self.handle = None
            self._scenes = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
