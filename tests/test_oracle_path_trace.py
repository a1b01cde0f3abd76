"""Schema v2 path trace records from the oracle (CPU).

The record restates include/psycles/luisa/path_trace_schema.h:10-121 as
written by src/luisa/path_kernel_setup.cpp:540-560 (global slots) and
src/luisa/path_kernel_surface_shading.cpp:270-340 (surface event slots 0-13
and 38).  Checks tie the slots to independent quantities: the sampler
probe's pixel hash, the discrete event log, the camera ray geometry and
the hit distance along each event's ray.
"""
import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, scenes

GLOBAL, EVENT, MAX_EVENTS = 8, 72, 4


def event(slots, i, slot):
    return slots[GLOBAL + i * EVENT + slot]


@pytest.fixture(scope="module")
def cornell():
    snap = scenes.cornell_box()
    return ob.OracleScene(snap), scenes.cornell_settings(32, 20)


@pytest.mark.parametrize("pixel, sample", [((16, 10), 0), ((3, 17), 5), ((30, 1), 11)])
def test_globals_match_hash_and_camera(cornell, pixel, sample):
    scene, s = cornell
    x, y = pixel
    slots, _ = scene.path_trace(s, x, y, sample, 16)
    assert slots.shape == (abi.PATH_TRACE_SLOTS, 4)
    assert list(slots[0, :3]) == [2.0, float(x), float(s.height - 1 - y)]
    probe = ob.sampler_probe(s, sample, 16, 0, 1)
    h = int(probe[y, x, 0])
    assert list(slots[1, :3]) == [float(sample), float(h & 0xFFFF), float(h >> 16)]
    fs = probe[y, x, 1:3].view(np.float32)
    expected = (0.5, 0.5) if sample == 0 else tuple(fs)
    assert tuple(slots[2, :2]) == expected
    assert np.all(slots[:7, 3] == 1.0) and slots[7, 3] == 0.0
    d = slots[5, :3]
    assert abs(np.linalg.norm(d) - 1.0) < 1e-6
    # perspective camera at (0, 0, 3.9) looking down -z: the ray starts at
    # the near clip along its own direction
    origin = np.array([0.0, 0.0, 3.9], np.float32)
    along = slots[4, :3] - origin
    assert np.allclose(np.cross(along, d), 0.0, atol=1e-6) and along @ d > 0.0


def test_events_agree_with_event_log_and_geometry(cornell):
    scene, s = cornell
    written = 0
    for y in range(0, 20, 3):
        for x in range(0, 32, 3):
            slots, records = scene.path_trace(s, x, y, 3, 16)
            log = scene.trace_events(s, 3, 16)[y, x]
            assert np.array_equal(records, log)
            for i in range(MAX_EVENTS):
                flags = int(records[i, abi.TRACE_FLAGS])
                if event(slots, i, 32)[3] == 1.0:
                    assert flags & 32               # post slots only after a sampled scatter
                hit = flags & 2 and not flags & 8
                w = event(slots, i, 0)[3]
                # slots are written exactly for surface events that survive the roulette
                assert (w == 1.0) == bool(hit and records[i, 7] != 0), (x, y, i, flags)
                if w != 1.0:
                    assert not np.any(slots[GLOBAL + i * EVENT:GLOBAL + (i + 1) * EVENT])
                    continue
                written += 1
                assert event(slots, i, 0)[0] == i
                assert event(slots, i, 0)[1] == records[i, 1]          # rng offset
                assert event(slots, i, 9)[0] == float(records[i, 2])   # object
                assert event(slots, i, 9)[1] == float(records[i, 3])   # primitive
                assert event(slots, i, 3)[2] == records[i, 7:8].view(np.float32)[0]  # RR probability
                o, d, t = event(slots, i, 5)[:3], event(slots, i, 6)[:3], event(slots, i, 8)[0]
                assert np.allclose(o + t * d, event(slots, i, 11)[:3], atol=1e-4)
                assert abs(np.linalg.norm(event(slots, i, 13)[:3]) - 1.0) < 1e-5
                assert event(slots, i, 7)[0] <= t <= event(slots, i, 7)[1]
                # sampler dimensions of this event's RNG offset: terminate,
                # light, light-terminate and BSDF samples
                dims = ob.sampler_probe(s, 3, 16, int(records[i, 1]), 4)[y, x, 1:].view(np.float32).reshape(4, 4)
                assert event(slots, i, 14)[0] == dims[0, 0] and event(slots, i, 14)[1] == dims[2, 0]
                assert np.array_equal(event(slots, i, 15)[:3], dims[1, :3])
                assert np.array_equal(event(slots, i, 16)[:3], dims[3, :3])
                post = event(slots, i, 32)
                if post[3] == 1.0 and i + 1 < MAX_EVENTS and event(slots, i + 1, 0)[3] == 1.0:
                    # the scattered ray and state are the next event's
                    assert np.array_equal(event(slots, i, 34)[:3], event(slots, i + 1, 5)[:3])
                    assert np.array_equal(event(slots, i, 35)[:3], event(slots, i + 1, 6)[:3])
                    assert post[2] == event(slots, i + 1, 0)[1] and post[0] == event(slots, i + 1, 0)[2]
                if i == 0:
                    assert np.array_equal(event(slots, 0, 5)[:3], slots[4, :3])
                    assert np.array_equal(event(slots, 0, 6)[:3], slots[5, :3])
                    assert list(event(slots, 0, 4)[:3]) == [1.0, 1.0, 1.0]
                else:
                    # the next ray leaves the previous surface point
                    assert np.allclose(event(slots, i, 5)[:3], event(slots, i - 1, 11)[:3], atol=1e-4)
    assert written > 100


def test_state_counters_follow_bounces():
    """Opaque diffuse walls: every event is one more diffuse bounce."""
    scene = ob.OracleScene(scenes.cornell_box())
    s = scenes.cornell_settings(8, 8)
    s.integrator.min_bounces = 8
    slots, _ = scene.path_trace(s, 4, 4, 1, 4)
    for i in range(MAX_EVENTS):
        if event(slots, i, 0)[3] != 1.0:
            break
        depth = event(slots, i, 0)
        lobes = event(slots, i, 1)
        assert depth[2] == i and lobes[0] == 0.0 and lobes[1] == i and lobes[2] == 0.0
        assert depth[1] == 16 * (i + 1)
        assert event(slots, i, 3)[2] == 1.0        # before the minimum bounce count


def test_decode_and_raw_round_trip(cornell, tmp_path):
    from psycles_amd import path_trace
    scene, s = cornell
    slots, records = scene.path_trace(s, 16, 10, 2, 16)
    trace = path_trace.PathTrace(16, s.height - 1 - 10, 2, slots)
    path_trace.write_raw_path_trace(trace, tmp_path / "trace.json")
    back = path_trace.read_raw_path_trace(tmp_path / "trace.json")
    assert np.array_equal(back.slots, slots) and (back.pixel_x, back.pixel_y, back.sample) == (16, 9, 2)
    decoded = path_trace.decode(back)
    assert decoded["global"]["header"] == (2.0, 16.0, 9.0)
    assert "reserved" not in decoded["global"]
    events = decoded["events"]
    assert 1 <= len(events) <= 4
    for i, e in enumerate(events):
        assert e["state_depth"][0] == i and e["isect_id"][0] == records[i, abi.TRACE_OBJECT]
        assert e["path_flag"] == path_trace.uint32_from_halves(*e["state_flags"][:2])
    # schema index helpers agree with the header's layout
    assert path_trace.index(0, "state_depth") == 8 and path_trace.index(3, "nee_contribution") == 8 + 3 * 72 + 47
    assert path_trace.closure_index(1, 7, 2) == 8 + 72 + 48 + 23
    assert path_trace.shadow_index(3, "shadow_transmittance") == 296 + 3 * 8 + 7 == 327


def shadow(slots, i, slot):
    return slots[296 + i * 8 + slot]


def traced_events(scene, s, pixels, sample=2, total=16):
    for x, y in pixels:
        slots, records = scene.path_trace(s, x, y, sample, total)
        for i in range(MAX_EVENTS):
            if event(slots, i, 0)[3] == 1.0:
                yield slots, records, i


@pytest.mark.parametrize("case", ["cornell", "cornell_lamps"])
def test_light_sample_and_shadow_slots(case):
    """The NEE record ties to the event log, the sampled point and the
    shadow ray: the shadow ray runs from its origin to the sampled light
    position (emissive triangles), its transmittance lies in [0, 1], and the
    contribution is the unshadowed estimate times that transmittance."""
    snap = scenes.cornell_box() if case == "cornell" else scenes.cornell_lamps()
    scene, s = ob.OracleScene(snap), scenes.cornell_settings(24, 16)
    sampled = shadows = contributions = 0
    for slots, records, i in traced_events(scene, s, [(x, y) for y in range(0, 16, 3) for x in range(0, 24, 3)]):
        meta = event(slots, i, 17)
        if meta[3] != 1.0 or meta[0] < 0.0:
            continue
        sampled += 1
        assert int(meta[1]) == records[i, abi.TRACE_LIGHT]        # the light the event log chose
        pdf = event(slots, i, 19)
        assert pdf[0] > 0.0 and 0.0 < pdf[1] <= 1.0
        d = event(slots, i, 20)[:3]
        assert abs(np.linalg.norm(d) - 1.0) < 1e-5
        nee_eval = event(slots, i, 23)
        if nee_eval[3] == 1.0:
            assert 0.0 <= nee_eval[2] <= 1.0                      # MIS weight
        sh = shadow(slots, i, 0)
        if sh[3] != 1.0:
            continue
        shadows += 1
        o, dr, rng = sh[:3], shadow(slots, i, 1)[:3], shadow(slots, i, 2)
        t = shadow(slots, i, 7)[:3]
        assert np.all((t >= 0.0) & (t <= 1.0))
        if meta[0] == 5.0:                                         # emissive triangle: ray ends on the sample
            assert np.allclose(o + dr * rng[1], event(slots, i, 21)[:3], atol=2e-4)
            assert shadow(slots, i, 4)[0] == event(slots, i, 18)[0]   # the light's own object is excluded
            assert shadow(slots, i, 4)[1] == meta[2]                  # and its primitive
        if np.any(t > 0.0):
            contributions += 1
            c = event(slots, i, 47)[:3]
            assert np.allclose(c, event(slots, i, 46)[:3] * t, rtol=1e-6, atol=0.0)
        else:
            assert event(slots, i, 47)[3] == 0.0
    assert sampled > 20 and shadows > 10 and contributions > 5


def test_closure_and_scatter_slots():
    """Closure slots follow the closure count of slot 10; the sampled
    closure is one of them, with its weight and normal; the BSDF pdf is the
    one post_mis keeps and the minimum never exceeds it."""
    snap = scenes.instanced_icospheres(frequency=4, instances=4, box=10.0)
    scene, s = ob.OracleScene(snap), scenes.instanced_settings(24, 16)
    picked = 0
    for slots, records, i in traced_events(scene, s, [(x, y) for y in range(0, 16, 2) for x in range(0, 24, 2)]):
        count = int(event(slots, i, 10)[2])
        for c in range(8):
            written = event(slots, i, 48 + 3 * c)[3] == 1.0
            assert written == (c < count)
            if written:
                assert event(slots, i, 48 + 3 * c)[0] == c and event(slots, i, 48 + 3 * c)[1] in (2.0, 3.0, 30.0)
        pick = event(slots, i, 24)
        if pick[3] != 1.0:
            continue
        picked += 1
        c = int(pick[0])
        assert c == records[i, abi.TRACE_CLOSURE] and c < count
        assert pick[1] == event(slots, i, 48 + 3 * c)[1] and pick[2] == event(slots, i, 48 + 3 * c)[2]
        assert np.array_equal(event(slots, i, 26), event(slots, i, 49 + 3 * c))
        assert 0.0 <= event(slots, i, 25)[2] < 1.0                # rescaled lobe random
        meta = event(slots, i, 28)
        assert meta[0] == meta[1]
        r = event(slots, i, 31)[:3]
        assert tuple(r) == ((1.0, 1.0, 1.0) if pick[1] in (2.0, 3.0) else (0.0, 0.0, 1.0))
        post = event(slots, i, 37)
        if post[3] == 1.0:
            assert post[1] <= post[0] or post[1] == np.float32(3.4028235e38)
            assert 0.0 < post[2] <= 1.0                           # continuation probability
            if pick[1] != 30.0:                                   # a non-transparent scatter sets the pdf
                assert post[0] == meta[0]
    assert picked > 30


def test_analytic_lamp_record():
    """Lamps carry Cycles' light type, no object and no shader id, and the
    light normal of their sampled point."""
    snap = scenes.cornell_lamps(with_quad=False)
    scene, s = ob.OracleScene(snap), scenes.cornell_settings(24, 16)
    types = set()
    for slots, records, i in traced_events(scene, s, [(x, y) for y in range(0, 16, 2) for x in range(0, 24, 2)]):
        meta = event(slots, i, 17)
        if meta[3] != 1.0 or meta[0] < 0.0:
            continue
        types.add(int(meta[0]))
        assert event(slots, i, 18)[0] == np.float32(4294967295.0)
        assert tuple(event(slots, i, 40)[:3]) == (65535.0, 65535.0, 0.0)
        n = event(slots, i, 22)[:3]
        assert abs(np.linalg.norm(n) - 1.0) < 1e-4
        sh = shadow(slots, i, 4)
        if sh[3] == 1.0:
            assert sh[0] == -1.0 and sh[1] == -1.0
    assert types and types <= {0, 1, 3, 4}
