"""Graph replay of wavefront dispatches, the launch-shape knobs and the
autotuner, through the C ABI.

A session captures its dispatch (counter reset, camera rays, every bounce
iteration's trace and shade, accumulation) into a graph once and replays it
per dispatch, with the dispatch's first sample read from device memory.  The
film after every replay, read back between replays with the host's own
copies and synchronisations in between, must be bit-identical to a session
that launches every kernel from the host.  Every launch shape in range
(tile width, refill lanes, trace chunk, shade chunks) must give the same
film bit for bit, and the autotuner must leave the session's film alone.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from psycles_amd import abi, scenes
from psycles_amd.render import RendererBackend, with_passes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_bit_identical(fa, fb, what=""):
    for name, x, y in zip(("combined", "normal", "albedo", "count"), fa, fb):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (what, name)


@pytest.fixture(scope="module")
def graphs():
    return RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=2)


@pytest.fixture(scope="module")
def eager():
    return RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=2, graphs=False)


def small_instanced():
    return scenes.instanced_icospheres(frequency=10, instances=12, box=14.0)


@pytest.mark.parametrize("name", ["cornell", "instanced"])
def test_graph_replays_match_host_launches(graphs, eager, name):
    if name == "cornell":
        snap, s = scenes.cornell_box(), scenes.cornell_settings(64, 40)
    else:
        snap, s = small_instanced(), scenes.instanced_settings(96, 54)
    log(f"case: graph replays vs host launches, {name}")
    ga = graphs.compile_scene(snap).scene
    ea = eager.compile_scene(snap).scene
    gs = graphs.create_session(ga, s)
    es = eager.create_session(ea, s)
    ranges = [(0, 2), (2, 4), (6, 1), (7, 5)]   # 1, 2, 1 and 3 dispatches of at most 2 samples
    for i, (first, count) in enumerate(ranges):
        gs.render_device(first, count, 12)
        es.render_device(first, count, 12)
        st = gs.stats()
        assert st.graph_launches == (count + 1) // 2, (i, st.graph_launches)
        assert eager_stats_ok(es.stats())
        assert_bit_identical(gs.read_film(), es.read_film(), (name, first, count))
        log(f"  range {first}+{count}: {st.graph_launches} replays, {st.graph_builds} graph builds, identical")
    # a parameter rebind changes the scene tables a graph was captured with
    mat = 2 if name == "cornell" else 0
    params = snap.materials[mat].parameters.copy()
    params[:3] = (0.9, 0.2, 0.7)
    assert ga.update_parameters(mat, params)
    assert ea.update_parameters(mat, params)
    gs.reset(s)
    es.reset(s)
    gs.render_device(0, 4, 4)
    es.render_device(0, 4, 4)
    assert gs.stats().graph_builds >= 1
    assert_bit_identical(gs.read_film(), es.read_film(), (name, "after rebind"))
    log("  after a parameter rebind: rebuilt and identical")


def eager_stats_ok(st) -> bool:
    return st.graph_launches == 0 and st.graph_builds == 0


def test_graph_reused_across_calls(graphs):
    snap, s = scenes.cornell_box(), scenes.cornell_settings(48, 30)
    log("case: one graph build serves every later range of the same size")
    sess = graphs.create_session(graphs.compile_scene(snap).scene, s)
    sess.render_device(0, 2, 8)
    assert sess.stats().graph_builds == 1
    for first in (2, 4, 6):
        sess.render_device(first, 2, 8)
        st = sess.stats()
        assert st.graph_builds == 0 and st.graph_launches == 1, (first, st.graph_builds, st.graph_launches)


def test_graph_with_light_passes(graphs, eager):
    light = ((abi.PASS_COMBINED, 4), (abi.PASS_DIFFUSE_DIRECT, 3), (abi.PASS_DIFFUSE_INDIRECT, 3),
             (abi.PASS_EMISSION, 3), (abi.PASS_ALBEDO, 3), (abi.PASS_SAMPLE_COUNT, 1))
    s, keep = with_passes(scenes.cornell_settings(40, 24), light)
    log("case: graph replays with light passes")
    snap = scenes.cornell_lamps()
    gs = graphs.create_session(graphs.compile_scene(snap).scene, s)
    es = eager.create_session(eager.compile_scene(snap).scene, s)
    for first, count in ((0, 3), (3, 3)):
        gs.render_device(first, count, 6)
        es.render_device(first, count, 6)
        assert_bit_identical(gs.read_film(), es.read_film(), ("light passes", first))
        assert np.array_equal(gs.read_light_passes().view(np.uint32), es.read_light_passes().view(np.uint32))
    assert keep is not None


SHAPES = [
    {"tile_width": 16}, {"tile_width": 32}, {"tile_width": 64},
    {"refill_lanes": 1}, {"refill_lanes": 64}, {"refill_lanes": 33},
    {"trace_chunk": 16}, {"trace_chunk": 4096}, {"trace_chunk": 200},
    {"queue_chunk": 128, "list_chunk": 64}, {"queue_chunk": 256, "list_chunk": 256},
    {"sort_shading": 1}, {"sort_shading": 1, "tile_width": 16, "list_chunk": 64},
    {"xcd_segments": 8}, {"xcd_segments": 8, "trace_chunk": 16, "refill_lanes": 1},
    {"tile_width": 64, "refill_lanes": 64, "trace_chunk": 16, "queue_chunk": 128, "list_chunk": 64},
]


@pytest.mark.parametrize("scene", ["cornell", "instanced"])
def test_every_launch_shape_gives_the_same_film(graphs, scene):
    if scene == "cornell":
        snap, s = scenes.cornell_box(), scenes.cornell_settings(64, 40)
        # a window whose sides are not multiples of any tile side
        s.window_x, s.window_y, s.window_width, s.window_height = 5, 3, 37, 23
    else:
        snap, s = small_instanced(), scenes.instanced_settings(96, 54)
    comp = graphs.compile_scene(snap).scene
    base = graphs.create_session(comp, s)
    base.render_device(0, 4, 4)
    ref = base.read_film()
    for shape in SHAPES:
        log(f"case: launch shape {shape} on {scene}")
        sess = graphs.create_session(comp, s)
        sess.set_tuning(shape)
        got = sess.tuning.as_dict()
        for k, v in shape.items():
            assert got[k] == v
        sess.render_device(0, 4, 4)
        assert_bit_identical(sess.read_film(), ref, (scene, shape))
        sess.close()


def test_invalid_shapes_rejected(graphs):
    s = scenes.cornell_settings(16, 16)
    sess = graphs.create_session(graphs.compile_scene(scenes.cornell_box()).scene, s)
    lib = graphs.lib
    default = abi.WavefrontTuning()
    lib.psx_wavefront_tuning_default(C.byref(default))
    assert default.as_dict() == {"tile_width": 8, "refill_lanes": 16, "trace_chunk": 128,
                                 "queue_chunk": 256, "list_chunk": 128, "sort_shading": 0, "xcd_segments": 1}
    for bad in ({"tile_width": 4}, {"tile_width": 12}, {"tile_width": 128}, {"refill_lanes": 0},
                {"refill_lanes": 65}, {"trace_chunk": 8}, {"trace_chunk": 8192}, {"queue_chunk": 64},
                {"queue_chunk": 512}, {"list_chunk": 32}, {"list_chunk": 512}, {"sort_shading": 2},
                {"xcd_segments": 0}, {"xcd_segments": 4}, {"xcd_segments": 16}):
        t = abi.WavefrontTuning.from_buffer_copy(default)
        for k, v in bad.items():
            setattr(t, k, v)
        assert lib.psx_session_set_tuning(sess.handle, C.byref(t)) == abi.PSX_ERROR_INVALID_ARGUMENT, bad
    assert sess.tuning.as_dict() == default.as_dict()


def test_autotune_keeps_the_film(graphs):
    snap, s = small_instanced(), scenes.instanced_settings(128, 72)
    comp = graphs.compile_scene(snap).scene
    ref = graphs.create_session(comp, s)
    ref.render_device(0, 2, 8)
    tuned = graphs.create_session(comp, s)
    tuned.render_device(0, 2, 8)
    log("case: autotune between two renders of a session")
    rep = tuned.autotune(2)
    log(f"  chosen {rep.chosen.as_dict()} {rep.chosen_ms:.3f} ms vs default {rep.default_ms:.3f} ms "
        f"over {rep.candidates} shapes")
    assert rep.candidates == 16 and rep.probe_samples == 2
    assert 0.0 < rep.chosen_ms <= rep.default_ms
    assert tuned.tuning.as_dict() == rep.chosen.as_dict()
    # the probes left the film alone; the next range renders with the new shape
    assert_bit_identical(tuned.read_film(), ref.read_film(), "after autotune")
    ref.render_device(2, 6, 8)
    tuned.render_device(2, 6, 8)
    assert_bit_identical(tuned.read_film(), ref.read_film(), "tuned render")


def test_autotune_before_first_render(graphs):
    snap, s = scenes.cornell_box(), scenes.cornell_settings(64, 40)
    comp = graphs.compile_scene(snap).scene
    sess = graphs.create_session(comp, s)
    rep = sess.autotune(1)
    assert rep.probe_samples == 1
    sess.render_device(0, 4, 16)
    ref = graphs.create_session(comp, s)
    ref.render_device(0, 4, 16)
    assert_bit_identical(sess.read_film(), ref.read_film(), "autotune first")


def test_autotune_megakernel_unsupported():
    mega = RendererBackend(0, integrator=abi.INTEGRATOR_MEGAKERNEL)
    sess = mega.create_session(mega.compile_scene(scenes.cornell_box()).scene, scenes.cornell_settings(16, 16))
    rep = abi.TuningReport()
    assert mega.lib.psx_session_autotune(sess.handle, 2, C.byref(rep)) == abi.PSX_ERROR_UNSUPPORTED
    assert mega.lib.psx_session_autotune(sess.handle, 0, C.byref(rep)) == abi.PSX_ERROR_INVALID_ARGUMENT


def test_timed_dispatches_launch_from_the_host(graphs):
    """Per-launch trace timing needs host launches; the film is unchanged."""
    timed = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=2, time_trace=True)
    snap, s = scenes.cornell_box(), scenes.cornell_settings(48, 30)
    a = timed.create_session(timed.compile_scene(snap).scene, s)
    b = graphs.create_session(graphs.compile_scene(snap).scene, s)
    a.render_device(0, 4, 4)
    b.render_device(0, 4, 4)
    sa, sb = a.stats(), b.stats()
    assert sa.trace_launches > 0 and sa.trace_ms > 0.0 and sa.graph_launches == 0
    assert sb.trace_launches == 0 and sb.graph_launches == 2
    assert_bit_identical(a.read_film(), b.read_film(), "timed vs graph")
