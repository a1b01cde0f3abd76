"""Wavefront integrator vs the megakernel and the CPU oracle, through the C ABI.

The wavefront path keeps path state in HBM between its trace and shade
kernels but runs the same path-step code and adds each pixel's samples in
sample order, so its film must be bit-identical to the megakernel's, and its
ray count equal.  Against the oracle the usual film bar applies (relative
error <= 1e-4 on >= 99.9 % of pixels).
"""
import sys

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.render import MemoryOutputSink, RendererBackend, SampleRange, with_passes

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4
PIXEL_FRACTION = 0.999


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mega():
    return RendererBackend(0, integrator=abi.INTEGRATOR_MEGAKERNEL)


@pytest.fixture(scope="module")
def wave():
    return RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)


def small_instanced():
    return scenes.instanced_icospheres(frequency=10, instances=12, box=14.0)


def render(backend, snap, s, first, count, total):
    comp = backend.compile_scene(snap)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    session = backend.create_session(comp.scene, s)
    session.render_device(first, count, total)
    return session.read_film(), session.stats()


def assert_bit_identical(fa, fb):
    for x, y in zip(fa, fb):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def relative_ok(a, b, tol=REL_TOL):
    diff = np.abs(a - b).max(axis=-1)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)).max(axis=-1), 1e-3)
    return diff <= tol * scale


def cornell_variants():
    out = [("default", scenes.cornell_settings(64, 40))]
    s = scenes.cornell_settings(64, 40)
    s.integrator.direct_light_sampling = abi.DIRECT_LIGHT_FORWARD
    out.append(("forward", s))
    s = scenes.cornell_settings(64, 40)
    s.integrator.direct_light_sampling = abi.DIRECT_LIGHT_NEE
    out.append(("nee", s))
    s = scenes.cornell_settings(64, 40)
    s.integrator.min_bounces = 0
    s.integrator.sample_clamp_indirect = 0.5
    s.integrator.sample_clamp_direct = 4.0
    out.append(("rr+clamp", s))
    s = scenes.cornell_settings(64, 40)
    s.transparent_background = 1
    s.pixel_filter = abi.FILTER_BLACKMAN_HARRIS
    s.filter_width = 1.5
    out.append(("filter+alpha", s))
    s = scenes.cornell_settings(96, 60)
    s.window_x, s.window_y, s.window_width, s.window_height = 16, 8, 41, 30
    out.append(("ragged window", s))
    return out


@pytest.mark.parametrize("name,settings", cornell_variants(), ids=[n for n, _ in cornell_variants()])
def test_wavefront_matches_megakernel_cornell(mega, wave, name, settings):
    log(f"case: wavefront vs megakernel, cornell {name}")
    snap = scenes.cornell_box()
    fm, sm = render(mega, snap, settings, 3, 8, 16)
    fw, sw = render(wave, snap, settings, 3, 8, 16)
    assert_bit_identical(fm, fw)
    assert sm.rays == sw.rays
    assert sw.trace_launches == 0 and sw.trace_ms == 0.0   # not timed unless asked


def test_wavefront_matches_megakernel_instanced(mega, wave):
    log("case: wavefront vs megakernel, small instanced 128x72x8")
    snap, s = small_instanced(), scenes.instanced_settings(128, 72)
    fm, sm = render(mega, snap, s, 0, 8, 8)
    fw, sw = render(wave, snap, s, 0, 8, 8)
    assert_bit_identical(fm, fw)
    assert sm.rays == sw.rays


def test_wavefront_film_parity_cornell_full(wave):
    """Cornell box at its full 640x400, 16 spp, against the oracle."""
    log("case: wavefront film parity cornell 640x400x16")
    snap = scenes.cornell_box()
    s = scenes.cornell_settings()
    (g_comb, g_norm, g_alb, g_cnt), _ = render(wave, snap, s, 0, 16, 16)
    film, _ = ob.OracleScene(snap).render(s, 0, 16, 16)
    assert np.array_equal(g_cnt, film.sample_count)
    for name, g, c in (("combined", g_comb, film.combined), ("normal", g_norm, film.normal),
                       ("albedo", g_alb, film.albedo)):
        frac = relative_ok(g, c).mean()
        log(f"  {name}: {frac:.6f}")
        assert frac >= PIXEL_FRACTION, name


def test_wavefront_batch_size_invariance():
    log("case: wavefront batch size invariance")
    snap = scenes.cornell_box()
    s = scenes.cornell_settings(48, 30)
    films = []
    for per in (1, 3, 16):
        be = RendererBackend(0, max_samples_per_dispatch=per, integrator=abi.INTEGRATOR_WAVEFRONT)
        f, st = render(be, snap, s, 0, 16, 16)
        assert st.dispatches == -(-16 // per)
        films.append(f)
    assert_bit_identical(films[0], films[1])
    assert_bit_identical(films[0], films[2])


def test_wavefront_progressive_and_sink(wave):
    log("case: wavefront progressive chunks and sink")
    snap = scenes.cornell_box()
    s, keep = with_passes(scenes.cornell_settings(64, 40))
    comp = wave.compile_scene(snap)
    a = wave.create_session(comp.scene, s)
    a.render_device(0, 16, 16)
    b = wave.create_session(comp.scene, s)
    b.render_device(0, 5, 16)
    b.render_device(5, 11, 16)
    assert_bit_identical(a.read_film(), b.read_film())
    sink = MemoryOutputSink()
    c = wave.create_session(comp.scene, s)
    assert c.render_samples(SampleRange(0, 8, 0, 8), sink)
    assert np.all(sink.passes[abi.PASS_SAMPLE_COUNT] == 8.0)


def test_wavefront_parameter_update(wave):
    log("case: wavefront parameter rebind")
    snap = scenes.cornell_box()
    s = scenes.cornell_settings(64, 40)
    comp = wave.compile_scene(snap)
    new = snap.materials[2].parameters.copy()
    new[:3] = (0.9, 0.2, 0.7)
    assert comp.scene.update_parameters(2, new)
    a = wave.create_session(comp.scene, s)
    a.render_device(0, 4, 4)
    snap2 = scenes.cornell_box()
    snap2.materials[2].parameters = new
    f2, _ = render(wave, snap2, s, 0, 4, 4)
    assert_bit_identical(a.read_film(), f2)


def test_traversal_counters_do_not_change_film():
    log("case: traversal counters")
    snap, s = small_instanced(), scenes.instanced_settings(96, 54)
    for integ in (abi.INTEGRATOR_MEGAKERNEL, abi.INTEGRATOR_WAVEFRONT):
        plain = RendererBackend(0, integrator=integ)
        counted = RendererBackend(0, integrator=integ, count_traversal=True)
        fa, sa = render(plain, snap, s, 0, 2, 2)
        fb, sb = render(counted, snap, s, 0, 2, 2)
        assert_bit_identical(fa, fb)
        assert sa.node_visits == 0 and sb.node_visits > sb.rays
        assert sb.triangle_tests > 0


def test_wavefront_full_instanced_scene_matches_megakernel(mega, wave):
    """The 64-instance, 1,003,520-triangle scene at 1920x1080, 2 spp."""
    log("case: wavefront vs megakernel, full instanced scene 2 spp")
    snap, s = scenes.instanced_icospheres(), scenes.instanced_settings()
    fm, sm = render(mega, snap, s, 0, 2, 64)
    fw, sw = render(wave, snap, s, 0, 2, 64)
    log(f"  megakernel {sm.kernel_ms:.1f} ms, wavefront {sw.kernel_ms:.1f} ms "
        f"(trace {sw.trace_ms:.1f} ms), rays {sw.rays}")
    assert_bit_identical(fm, fw)
    assert sm.rays == sw.rays
