"""Render server on the GPU, through the C ABI.

Every request's film must be bit-identical to a standalone wavefront session
rendering the same sample range: alone, mixed with requests of other scenes,
resolutions, windows and sample ranges, across parameter edits submitted
between requests, and under an HBM budget that forces requests to queue.
"""
import sys

import numpy as np
import pytest

from psycles_amd import abi, scenes
from psycles_amd.render import MemoryOutputSink, RendererBackend, SampleRange, with_passes
from psycles_amd.server import RenderServer

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def backend():
    return RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=4)


def standalone(backend, scene, settings, first, count, total):
    sess = backend.create_session(scene, settings)
    sess.render_device(first, count, total)
    return sess.read_film()


def assert_same(a, b):
    for x, y in zip(a, b):
        assert x.shape == y.shape
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def small_instanced():
    return scenes.instanced_icospheres(frequency=12, instances=8)


def test_single_request_matches_session(backend):
    log("case: one request vs session (cornell 64x40, 8 spp)")
    comp = backend.compile_scene(scenes.cornell_box())
    s = scenes.cornell_settings(64, 40)
    server = RenderServer(backend, path_capacity=1 << 16)
    r = server.submit(comp.scene, s, 0, 8, 8)
    assert r.wait(60000)
    st = r.status()
    assert st.state == abi.REQUEST_DONE and st.samples_done == 8
    assert st.done_ms >= st.first_launch_ms >= st.admit_ms >= st.submit_ms
    assert_same(r.read_film(), standalone(backend, comp.scene, s, 0, 8, 8))
    server.close()


def test_mixed_requests_match_sessions(backend):
    log("case: 14 mixed requests over four scenes")
    snaps = {"cornell": scenes.cornell_box(), "lamps": scenes.cornell_lamps(), "instanced": small_instanced(),
             "many": scenes.many_lights(frequency=6, instances=6, emissive=150, points=12, spots=4)}
    comps = {k: backend.compile_scene(v).scene for k, v in snaps.items()}

    def settings(kind, w, h, seed, window=None):
        if kind in ("cornell", "lamps"):
            s = scenes.cornell_settings(w, h)
        elif kind == "instanced":
            s = scenes.instanced_settings(w, h)
        else:
            s = scenes.many_light_settings(w, h)
        s.seed = seed
        if window:
            s.window_x, s.window_y, s.window_width, s.window_height = window
        return s

    plan = [("cornell", 64, 40, 1, None, 0, 4, 4), ("instanced", 96, 54, 2, None, 0, 2, 8),
            ("lamps", 50, 30, 3, None, 0, 8, 8), ("many", 64, 40, 4, None, 0, 4, 4),
            ("cornell", 640, 400, 5, (100, 60, 77, 45), 3, 5, 16), ("instanced", 33, 17, 6, None, 0, 16, 16),
            ("cornell", 256, 256, 7, None, 0, 1, 1), ("lamps", 40, 40, 8, None, 2, 3, 8),
            ("many", 80, 48, 9, None, 0, 2, 32), ("instanced", 128, 72, 10, None, 4, 4, 64),
            ("cornell", 31, 29, 11, None, 0, 12, 12), ("many", 24, 24, 12, None, 1, 6, 8),
            ("lamps", 72, 40, 13, None, 0, 1, 4), ("cornell", 64, 40, 14, None, 8, 8, 16)]
    server = RenderServer(backend, path_capacity=1 << 17)
    reqs = []
    for kind, w, h, seed, window, first, count, total in plan:
        s = settings(kind, w, h, seed, window)
        reqs.append((kind, s, first, count, total, server.submit(comps[kind], s, first, count, total)))
    for kind, s, first, count, total, r in reqs:
        assert r.wait(120000), kind
    st = server.stats()
    log(f"  batches {st.batches}, paths {st.paths}, rays {st.rays}, peak HBM {st.peak_hbm}")
    assert st.completed == len(plan)
    for kind, s, first, count, total, r in reqs:
        log(f"  {kind} {s.width}x{s.height} range {first}+{count}/{total}: batches {r.status().batches}")
        assert_same(r.read_film(), standalone(backend, comps[kind], s, first, count, total))
    assert st.batches < sum(r.status().batches for *_, r in reqs), "no batch was shared"
    server.close()


def test_edits_apply_in_submission_order(backend):
    log("case: parameter edit between submissions")
    snap = scenes.cornell_box()
    comp = backend.compile_scene(snap).scene
    s = scenes.cornell_settings(64, 40)
    light = next(i for i, m in enumerate(snap.materials) if m.name == "light")
    wall = 2
    new_light = snap.materials[light].parameters.copy()
    new_light[-1] *= 0.5
    new_wall = snap.materials[wall].parameters.copy()
    new_wall[:3] = (0.2, 0.7, 0.9)
    server = RenderServer(backend, path_capacity=1 << 15)
    before = [server.submit(comp, s, 0, 16, 16) for _ in range(2)]
    server.update_parameters(comp, light, new_light)
    server.update_parameters(comp, wall, new_wall)
    after = server.submit(comp, s, 0, 16, 16)
    for r in before + [after]:
        assert r.wait(60000)
    old = standalone(backend, backend.compile_scene(scenes.cornell_box()).scene, s, 0, 16, 16)
    edited = scenes.cornell_box()
    edited.materials[light].parameters = new_light
    edited.materials[wall].parameters = new_wall
    new = standalone(backend, backend.compile_scene(edited).scene, s, 0, 16, 16)
    for r in before:
        assert_same(r.read_film(), old)
    assert_same(after.read_film(), new)
    assert server.stats().edits == 2
    server.close()


def test_budget_queues_requests(backend):
    log("case: HBM budget admits one film at a time")
    comp = backend.compile_scene(scenes.cornell_box()).scene
    s = scenes.cornell_settings(320, 200)
    film = 320 * 200 * 52 + 4096
    probe = RenderServer(backend, path_capacity=1 << 16, streams=2)
    fixed = probe.stats().hbm_in_use
    probe.close()
    server = RenderServer(backend, path_capacity=1 << 16, streams=2, hbm_budget=fixed + film + film // 2)
    reqs = [server.submit(comp, s, 0, 2, 2) for _ in range(4)]
    for r in reqs:
        assert r.wait(60000)
        assert_same(r.read_film(), standalone(backend, comp, s, 0, 2, 2))
        r.release()
    st = server.stats()
    assert st.peak_hbm <= fixed + film + film // 2
    server.close()


def test_server_beside_graph_captures(backend):
    log("case: server admissions while the caller's sessions capture dispatch graphs")
    # HIP refuses allocations and event queries on every thread in global
    # capture mode while any stream captures; the worker must not be one
    comp = backend.compile_scene(scenes.cornell_box()).scene
    s = scenes.cornell_settings(160, 100)
    server = RenderServer(backend, path_capacity=1 << 15, streams=2)
    reference = standalone(backend, comp, s, 0, 2, 2)
    reqs = []
    for _ in range(24):
        reqs.append(server.submit(comp, s, 0, 2, 2))
        sess = backend.create_session(comp, s)
        sess.render_device(0, 2, 2)
        assert sess.stats().graph_builds >= 1
        sess.close()
    for r in reqs:
        assert r.wait(60000)
        assert_same(r.read_film(), reference)
        r.release()
    server.close()


def test_cancel_and_validation(backend):
    log("case: cancel and invalid submissions")
    comp = backend.compile_scene(small_instanced()).scene
    s = scenes.instanced_settings(256, 144)
    server = RenderServer(backend, path_capacity=1 << 16, streams=2)
    big = server.submit(comp, s, 0, 64, 64)
    big.cancel()
    assert not big.wait(60000)
    assert big.status().state == abi.REQUEST_CANCELLED
    with pytest.raises(abi.PsxError):
        server.submit(comp, s, 4, 8, 8)            # range past total
    with pytest.raises(abi.PsxError):
        server.submit(comp, scenes.instanced_settings(1920, 1080), 0, 1, 1)  # window above the capacity
    ok = server.submit(comp, scenes.instanced_settings(64, 36), 0, 2, 2)
    assert ok.wait(60000)
    server.close()


def test_write_passes_match_session(backend):
    log("case: passes through a sink")
    comp = backend.compile_scene(scenes.cornell_box()).scene
    s, keep = with_passes(scenes.cornell_settings(48, 32))
    server = RenderServer(backend, path_capacity=1 << 15)
    r = server.submit(comp, s, 0, 4, 4)
    assert r.wait(60000)
    a = MemoryOutputSink()
    r.write_passes(a)
    b = MemoryOutputSink()
    sess = backend.create_session(comp, s)
    assert sess.render_samples(SampleRange(0, 4, 0, 4), b)
    assert len(a.passes) == 4 and sorted(a.passes) == sorted(b.passes)
    for k in a.passes:
        assert np.array_equal(a.passes[k], b.passes[k])
    server.close()
