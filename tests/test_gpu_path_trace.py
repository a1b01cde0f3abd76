"""Schema v2 path trace (psx_session_path_trace) on the GPU vs the oracle.

One pixel and sample per call, as LuisaPathTraceRequest does.  Two groups of
slots with their own bars, each on the traced pixels:

* the path itself: header, hash, state counters, visibility and flag words,
  hit identity, material / closure count, runtime flags, sampler values,
  closures, closure pick and BSDF sample, post-scatter state and MIS, and
  which of these slots are written, identical on >= 99.9 %; their float
  values (rays, throughput, hit distance and barycentrics, surface frame,
  closure weights, BSDF sample and pdfs) within 1e-4 relative (1e-5
  absolute floor) on >= 99.9 %;
* the light-sample record of NEE (slots 17-23 and 40-47: light type,
  emitter, primitive, object, shader, pdfs, sampled direction, point and
  normal, evaluation, NEE terms) and the shadow events (296-327): discrete
  words and written slots identical, floats within 1e-3 relative (1e-4
  absolute floor), each on >= 99.9 %, over the events whose light sample
  is not grazing (|cos| >= 0.02 between the sampled direction and the
  light's normal, on both sides).  The emissive-triangle sampler inverts a
  spherical triangle (acosf, atan2f), which is ill-conditioned where the
  shading point lies almost in the emitter's plane: there the oracle's glibc
  and the device's ocml transcendentals move the sampled point by up to a
  few per cent and can flip the sample's side test.  Those events (at most
  15 % of the sampled ones; 7.9 % in the Cornell box, whose shading points
  near the ceiling see its emitter edge-on) are counted and left out; the single-function
  probes (test_gpu_fixtures.py) hold the sampler to the oracle bit for bit
  where it is well conditioned.

The discrete words of the same call's event log (flags, RNG offset, hit,
light and closure choice) equal the oracle's on >= 99.9 %.
"""
import sys

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, path_trace, scenes
from psycles_amd.render import RendererBackend

pytestmark = pytest.mark.gpu

GLOBAL, EVENT = 8, 72
# state counters, hit identity, material, sampler values, roughness / eta of
# the BSDF sample, post-scatter state and flag words
DISCRETE_EVENT = (0, 1, 2, 9, 10, 14, 15, 16, 31, 32, 36, 38, 39)
# the light-sample record and its NEE terms
NEE_EVENT = tuple(range(17, 24)) + tuple(range(40, 48))
NEE_DISCRETE = (17, 18, 40, 45)
# shadow-event slots: source, light and first-hit identities
DISCRETE_SHADOW = (3, 4, 5)
PIXEL_FRACTION = 0.999
GRAZING = 0.02


def grazing_events(slots) -> list:
    """Events whose light sample meets the light at |cos| < GRAZING."""
    out = []
    for e in range(4):
        base = GLOBAL + e * EVENT
        d, n = slots[base + 20], slots[base + 22]
        if d[3] == 1.0 and n[3] == 1.0 and abs(float(np.dot(d[:3], n[:3]))) < GRAZING:
            out.append(e)
    return out


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def nee_rows(events=range(4)):
    """Slot rows of the light-sample record and the shadow events."""
    rows = np.zeros(abi.PATH_TRACE_SLOTS, bool)
    for e in events:
        base = GLOBAL + e * EVENT
        rows[[base + s for s in NEE_EVENT]] = True
        rows[296 + e * 8: 296 + (e + 1) * 8] = True
    return rows


def discrete_mask():
    """Exactly compared entries: discrete words and every written flag."""
    m = np.zeros((abi.PATH_TRACE_SLOTS, 4), bool)
    m[[0, 1, 3], :] = True
    m[:, 3] = True                               # written flags
    for e in range(4):
        base = GLOBAL + e * EVENT
        m[[base + s for s in DISCRETE_EVENT + NEE_DISCRETE], :] = True
        m[base + 3, :2] = True                   # path flag halves; [2] is the RR probability
        m[[296 + e * 8 + s for s in DISCRETE_SHADOW], :] = True
    return m


def path_trace_slot_name(i: int) -> str:
    if i < GLOBAL:
        return path_trace.GLOBAL_SLOTS[i]
    if i >= path_trace.SHADOW_EVENT_BASE:
        return "shadow." + path_trace.SHADOW_EVENT_SLOTS[(i - path_trace.SHADOW_EVENT_BASE) % 8]
    j = (i - GLOBAL) % EVENT
    return path_trace.EVENT_SLOTS[j] if j < len(path_trace.EVENT_SLOTS) else f"closure.{(j - 48) % 3}"


def close(a, b, rel=1.0e-4, floor=1.0e-5):
    return np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)) + floor


@pytest.mark.parametrize("case", ["cornell_lamps", "instanced", "smooth_mixed", "many_lights"])
def test_path_trace_gpu_vs_oracle(case):
    log(f"path trace {case}")
    if case == "cornell_lamps":
        snap, s = scenes.cornell_lamps(), scenes.cornell_settings(40, 25)
    elif case == "instanced":
        snap = scenes.instanced_icospheres(frequency=6, instances=6, box=12.0)
        s = scenes.instanced_settings(40, 24)
    elif case == "many_lights":
        # light tree, emissive triangles and 17 lamps (forward hits through the lamp BVH)
        snap = scenes.many_lights(frequency=6, instances=6, emissive=150, points=12, spots=4)
        s = scenes.many_light_settings(40, 24)
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
    nee = nee_rows()
    core = ~nee[:, None]
    counts = {"core_discrete": 0, "core_float": 0, "nee_discrete": 0, "nee_float": 0, "events": 0}
    total = written = sampled = grazing = 0
    differing, float_slots, float_examples = [], {}, []
    for sample in (0, 5):
        for y in range(0, s.height, 2):
            for x in range(0, s.width, 2):
                # the request names the row counted from the lower edge of the image
                g_slots, g_rec = session.path_trace(x, s.height - 1 - y, sample, 8)
                o_slots, o_rec = oracle.path_trace(s, x, y, sample, 8)
                total += 1
                written += int(g_slots[GLOBAL:, 3].sum())
                equal = g_slots == o_slots
                sampled += sum(1 for e in range(4) if o_slots[GLOBAL + e * EVENT + 17, 3] == 1.0)
                skip = sorted(set(grazing_events(g_slots)) | set(grazing_events(o_slots)))
                grazing += len(skip)
                compared = ~core & ~nee_rows(skip)[:, None]
                for group, rows in (("core", core), ("nee", compared)):
                    m = mask & rows
                    same = bool(np.all(equal[m]))
                    counts[group + "_discrete"] += same
                    if not same and len(differing) < 8:
                        bad = np.nonzero(np.any(~equal & m, axis=1))[0]
                        differing.append((x, y, sample, [(path_trace_slot_name(int(r)), g_slots[r].tolist(),
                                                          o_slots[r].tolist()) for r in bad[:3]]))
                near_core = close(g_slots, o_slots) | mask | ~core
                near_nee = close(g_slots, o_slots, 1.0e-3, 1.0e-4) | mask | ~compared
                counts["core_float"] += bool(near_core.all())
                counts["nee_float"] += bool(near_nee.all())
                for r in np.nonzero(~np.all(near_core & near_nee, axis=1))[0]:
                    name = path_trace_slot_name(int(r))
                    float_slots[name] = float_slots.get(name, 0) + 1
                    if len(float_examples) < 8:
                        float_examples.append((x, y, sample, name, g_slots[r].tolist(), o_slots[r].tolist()))
                counts["events"] += bool(np.array_equal(g_rec[:, :6], o_rec[:, :6]))
    log(f"  {total} traces, {written} event slots written, {sampled} light samples ({grazing} grazing, "
        f"left out): " + ", ".join(f"{k} {v / total:.4f}" for k, v in counts.items()))
    for d in differing:
        log(f"  discrete slots differ at pixel {d[0]},{d[1]} sample {d[2]}: {d[3]}")
    if float_slots:
        log(f"  traces with a float slot outside its bar, by slot: {float_slots}")
        for e in float_examples:
            log(f"  {e[3]} at pixel {e[0]},{e[1]} sample {e[2]}: device {e[4]} oracle {e[5]}")
    assert written > 0
    assert counts["core_discrete"] / total >= PIXEL_FRACTION
    assert counts["core_float"] / total >= PIXEL_FRACTION
    assert counts["events"] / total >= PIXEL_FRACTION
    assert counts["nee_discrete"] / total >= PIXEL_FRACTION
    assert counts["nee_float"] / total >= PIXEL_FRACTION
    assert grazing <= 0.15 * max(sampled, 1)    # 56 of 709 (7.9 %) in the Cornell lamps case
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
