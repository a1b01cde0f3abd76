"""The serving request stream (host logic, no GPU)."""
import numpy as np

from psycles_amd import abi, scenes, serving


def snapshots():
    return {"cornell": scenes.cornell_box(), "lamps": scenes.cornell_lamps()}


def test_stream_is_deterministic_and_shaped():
    a = serving.make_stream(snapshots(), 400)
    b = serving.make_stream(snapshots(), 400)
    assert [(r.arrival, r.scene, r.edit, r.width, r.spp) for r in a] == \
           [(r.arrival, r.scene, r.edit, r.width, r.spp) for r in b]
    arrivals = np.array([r.arrival for r in a])
    assert np.all(np.diff(arrivals) > 0)
    # exponential inter-arrival at 200 req/s: mean 5 ms
    assert 0.004 < np.diff(arrivals).mean() < 0.006
    edits = sum(r.edit for r in a)
    assert 0.24 < edits / len(a) < 0.36
    for r in a:
        if not r.edit:
            assert (r.width, r.height) in serving.RESOLUTIONS and r.spp in serving.SAMPLE_COUNTS
            assert r.viewport == (r.width * r.height <= 256000 and r.spp <= 4)


def test_edits_scale_albedo_or_strength_and_chain():
    snaps = snapshots()
    original = {n: [m.parameters.copy() for m in s.materials] for n, s in snaps.items()}
    stream = serving.make_stream(snaps, 300)
    current = {n: [p.copy() for p in v] for n, v in original.items()}
    for r in (x for x in stream if x.edit):
        m = snaps[r.scene].materials[r.material]
        before = current[r.scene][r.material]
        after = r.parameters
        if any(i[0] == abi.CLOSURE_EMISSION for i in m.instructions):
            k = next(i for i in m.instructions if i[0] == abi.CLOSURE_EMISSION)[5]
            ratio = after[k] / before[k]
            assert 0.8 - 1e-6 <= ratio <= 1.2 + 1e-6
        else:
            c = next(i for i in m.instructions if i[0] == abi.CLOSURE_DIFFUSE)[4]
            assert np.all(after[c:c + 3] <= 1.0)
        current[r.scene][r.material] = after
    # the snapshots are left as they were
    for n, s in snaps.items():
        for m, p in zip(s.materials, original[n]):
            assert np.array_equal(m.parameters, p)
