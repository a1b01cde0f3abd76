"""Schema v2 path trace (psx_session_path_trace) on the GPU vs the oracle.

One pixel and sample per call, as LuisaPathTraceRequest does.  Bars: the
discrete slots (header, hash, state counters, visibility and flag words,
hit identity, material / closure count, runtime flags, sampler values,
post-scatter state) and which slots are written are identical; the float slots (rays, throughput, hit distance,
barycentrics, surface frame, scattered ray and throughput) agree within 1e-4 relative (1e-5 absolute
floor), each on >= 99.9 % of the traced pixels; the discrete words of the
same call's event log (flags, RNG offset, hit, light and closure choice)
equal the oracle's.
"""
import sys

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, path_trace, scenes
from psycles_amd.render import RendererBackend

pytestmark = pytest.mark.gpu

GLOBAL, EVENT = 8, 72
# state counters, hit identity, material, sampler values, post-scatter state and flag words
DISCRETE_EVENT = (0, 1, 2, 9, 10, 14, 15, 16, 32, 36, 38, 39)
# throughput, rays, range, hit distance and barycentrics, surface frame
FLOAT_EVENT = (4, 5, 6, 7, 8, 11, 12, 13, 33, 34, 35)
PIXEL_FRACTION = 0.999


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def discrete_mask():
    m = np.zeros((abi.PATH_TRACE_SLOTS, 4), bool)
    m[[0, 1, 3], :] = True
    m[:, 3] = True                               # written flags
    for e in range(4):
        base = GLOBAL + e * EVENT
        m[[base + s for s in DISCRETE_EVENT], :] = True
        m[base + 3, :2] = True                   # path flag halves; [2] is the RR probability
    return m


def close(a, b):
    return np.abs(a - b) <= 1.0e-4 * np.maximum(np.abs(a), np.abs(b)) + 1.0e-5


@pytest.mark.parametrize("case", ["cornell_lamps", "instanced", "smooth_mixed"])
def test_path_trace_gpu_vs_oracle(case):
    log(f"path trace {case}")
    if case == "cornell_lamps":
        snap, s = scenes.cornell_lamps(), scenes.cornell_settings(40, 25)
    elif case == "instanced":
        snap = scenes.instanced_icospheres(frequency=6, instances=6, box=12.0)
        s = scenes.instanced_settings(40, 24)
    else:
        sys.path.insert(0, __file__.rsplit("/", 1)[0])
        import blender_bundle
        snap = blender_bundle.smooth_sphere_snapshot()
        s = abi.default_settings(40, 30)
    backend = RendererBackend()
    comp = backend.compile_scene(snap)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    session = backend.create_session(comp.scene, s)
    oracle = ob.OracleScene(snap)
    mask = discrete_mask()
    discrete_ok = float_ok = events_ok = total = written = 0
    for sample in (0, 5):
        for y in range(0, s.height, 2):
            for x in range(0, s.width, 2):
                # the request names the row counted from the lower edge of the image
                g_slots, g_rec = session.path_trace(x, s.height - 1 - y, sample, 8)
                o_slots, o_rec = oracle.path_trace(s, x, y, sample, 8)
                total += 1
                written += int(g_slots[GLOBAL:, 3].sum())
                discrete_ok += bool(np.array_equal(g_slots[mask], o_slots[mask]))
                float_ok += bool(close(g_slots[~mask], o_slots[~mask]).all())
                events_ok += bool(np.array_equal(g_rec[:, :6], o_rec[:, :6]))
    log(f"  {total} traces, {written} event slots written: discrete {discrete_ok / total:.4f} "
        f"float {float_ok / total:.4f} events {events_ok / total:.4f}")
    assert written > 0
    assert discrete_ok / total >= PIXEL_FRACTION
    assert float_ok / total >= PIXEL_FRACTION
    assert events_ok / total >= PIXEL_FRACTION
    session.close()
    comp.scene.close()
    backend.close()


def test_path_trace_rejects_out_of_window_pixel():
    snap, s = scenes.cornell_box(), scenes.cornell_settings(16, 8)
    backend = RendererBackend()
    comp = backend.compile_scene(snap)
    session = backend.create_session(comp.scene, s)
    for pixel_x, pixel_y, sample in ((16, 0, 0), (0, 8, 0), (0, 0, 4)):
        with pytest.raises(abi.PsxError):
            session.path_trace(pixel_x, pixel_y, sample, 4)
    # a window: pixels outside it are refused, inside it the record names the full-image pixel
    s2 = scenes.cornell_settings(16, 8)
    s2.window_x, s2.window_y, s2.window_width, s2.window_height = 4, 2, 8, 4
    windowed = backend.create_session(comp.scene, s2)
    with pytest.raises(abi.PsxError):
        windowed.path_trace(2, 3, 0, 4)
    trace = path_trace.trace_path(windowed, path_trace.PathTraceRequest(5, 3, 1), 4)
    assert list(trace.slots[0, :3]) == [2.0, 5.0, 3.0]
    o_slots, _ = ob.OracleScene(snap).path_trace(s2, 1, 2, 1, 4)
    assert np.array_equal(trace.slots[0], o_slots[0])
    windowed.close()
    session.close()
    comp.scene.close()
    backend.close()
