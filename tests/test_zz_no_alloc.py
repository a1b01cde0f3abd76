"""psx_scene_intersect_device_noalloc (static scratch, no allocation) against
psx_scene_intersect_device (allocating) through the C ABI: the same hit
records, byte for byte, on two calls in a row, the second smaller; and the
size limit rejected before any launch."""
import ctypes as C
import sys

import numpy as np
import pytest

from psycles_amd import abi, scenes
from psycles_amd.render import RendererBackend

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class DeviceArray:
    """hipMalloc'd bytes freed on close (the test's own buffers)."""

    hip = None

    def __init__(self, nbytes):
        if DeviceArray.hip is None:
            DeviceArray.hip = C.CDLL("libamdhip64.so")
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        assert DeviceArray.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(nbytes, 16))) == 0

    def upload(self, host):
        assert DeviceArray.hip.hipMemcpy(self.ptr, C.cast(host, C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def download(self, host):
        assert DeviceArray.hip.hipMemcpy(C.cast(host, C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0

    def close(self):
        if self.ptr:
            DeviceArray.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def random_queries(n, rng):
    o = rng.uniform(-15.0, 15.0, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    kinds = rng.integers(0, 2, n)
    recs = []
    for i in range(n):
        shadow = kinds[i] == abi.QUERY_SHADOW
        recs.append({"origin": o[i], "direction": d[i], "tmin": 0.0, "tmax": float(rng.uniform(1.0, 40.0)),
                     "visibility": abi.VIS_SHADOW if shadow else abi.VIS_ALL, "kind": int(kinds[i]),
                     "transparent_max": 8})
    return abi.make_queries(recs)


def run(lib, scene, queries, entry):
    n = len(queries)
    dq, dh = DeviceArray(n * 64), DeviceArray(n * 48)
    try:
        dq.upload(queries)
        rc = getattr(lib, entry)(scene.handle, dq.ptr, n, dh.ptr)
        assert rc == abi.PSX_OK, (entry, rc)
        hits = (abi.RayHit * n)()
        dh.download(hits)
        return bytes(hits)
    finally:
        dq.close()
        dh.close()


def test_noalloc_matches_allocating_entry_point():
    backend = RendererBackend(0)
    comp = backend.compile_scene(scenes.instanced_icospheres(frequency=10, instances=12, box=14.0))
    assert comp.ok()
    lib = backend.lib
    rng = np.random.default_rng(21)
    for n in (50000, 777):
        log(f"case: no-alloc ray queries, {n} queries")
        q = random_queries(n, rng)
        ref = run(lib, comp.scene, q, "psx_scene_intersect_device")
        got = run(lib, comp.scene, q, "psx_scene_intersect_device_noalloc")
        assert got == ref
        hits = np.frombuffer(got, dtype=np.uint32).reshape(n, 12)
        log(f"  identical; {int(hits[:, 5].sum())} hit or occluded")
    log("case: no-alloc size limit")
    dq, dh = DeviceArray(64), DeviceArray(48)
    try:
        rc = lib.psx_scene_intersect_device_noalloc(comp.scene.handle, dq.ptr, abi.INTERSECT_NOALLOC_MAX_QUERIES + 1,
                                                    dh.ptr)
        assert rc == abi.PSX_ERROR_INVALID_ARGUMENT
    finally:
        dq.close()
        dh.close()
