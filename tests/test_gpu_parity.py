"""HIP render path vs the CPU oracle, through the C ABI.

Bars (north-star parity gates):
* Sobol/RNG dimension values: bit-exact.
* primary-hit identities: identical (ties resolve by the same total order).
* discrete path events (light choice, closure choice, RR, hit identity):
  bit-exact on every recorded event (each flip would be counted by kind and
  attributed to the first path-trace slot that differs; none occurs).
* Combined / Normal / Albedo: relative error <= 1e-4 on >= 99.9 % of pixels.
"""
import math
import sys

import numpy as np
import pytest

import event_flips
import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.render import (MemoryOutputSink, RendererBackend, SampleRange,
                                with_passes)

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4
PIXEL_FRACTION = 0.999


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def backend():
    return RendererBackend(0)


def compile_scene(backend, snap):
    comp = backend.compile_scene(snap)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    return comp.scene


def relative_ok(a, b, tol=REL_TOL):
    """Per-pixel relative agreement: |a-b| <= tol * max(|a|, |b|, 1e-3)."""
    diff = np.abs(a - b).max(axis=-1)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)).max(axis=-1), 1e-3)
    return diff <= tol * scale


def small_instanced():
    return scenes.instanced_icospheres(frequency=10, instances=12, box=14.0)


# ------------------------------------------------------------------ sampler

def test_sampler_dimensions_bit_exact(backend):
    log("case: sampler probe cornell 64x40")
    snap = scenes.cornell_box()
    scene = compile_scene(backend, snap)
    s = scenes.cornell_settings(64, 40)
    session = backend.create_session(scene, s)
    for sample in (0, 5, 15):
        gpu = session.sampler_probe(sample, 16, 0, 36)
        cpu = ob.sampler_probe(s, sample, 16, 0, 36)
        assert np.array_equal(gpu, cpu), sample


# ------------------------------------------------------------------ primary hits

@pytest.mark.parametrize("name", ["cornell", "instanced"])
def test_primary_hits_identical(backend, name):
    log(f"case: primary hits {name}")
    if name == "cornell":
        snap, s = scenes.cornell_box(), scenes.cornell_settings(96, 60)
    else:
        snap, s = small_instanced(), scenes.instanced_settings(96, 54)
    scene = compile_scene(backend, snap)
    session = backend.create_session(scene, s)
    osc = ob.OracleScene(snap)
    for sample in (0, 3):
        gpu = session.primary_hits(sample, 8)
        cpu = osc.primary_hits(s, sample, 8)
        assert np.array_equal(gpu, cpu)


# ------------------------------------------------------------------ events

@pytest.mark.parametrize("name", ["cornell", "instanced"])
def test_path_events(backend, name):
    log(f"case: path events {name}")
    if name == "cornell":
        snap, s = scenes.cornell_box(), scenes.cornell_settings(80, 50)
    else:
        snap, s = small_instanced(), scenes.instanced_settings(80, 45)
    scene = compile_scene(backend, snap)
    session = backend.create_session(scene, s)
    osc = ob.OracleScene(snap)
    event_flips.check(f"path events {name}", session, osc, s, (0, 7), 16, log)
    for sample in (0, 7):
        gpu = session.trace_events(sample, 16)
        cpu = osc.trace_events(s, sample, 16)
        valid = (cpu[..., 0] & 1) != 0
        # the continuation probability is a float of the throughput chain:
        # it may differ in the last ulp after a libm/ocml sin or cos
        pg = gpu[..., abi.TRACE_RR_BITS][valid].view(np.float32)
        pc = cpu[..., abi.TRACE_RR_BITS][valid].view(np.float32)
        assert np.all(np.abs(pg - pc) <= 1e-6 * np.maximum(np.abs(pc), 1e-30))


# ------------------------------------------------------------------ film

def film_parity(backend, snap, s, first, count, total, threads=0):
    scene = compile_scene(backend, snap)
    session = backend.create_session(scene, s)
    session.render_device(first, count, total)
    g_comb, g_norm, g_alb, g_cnt = session.read_film()
    film, _ = ob.OracleScene(snap).render(s, first, count, total, threads=threads)
    assert np.array_equal(g_cnt, film.sample_count)
    results = {}
    for name, g, c in (("combined", g_comb, film.combined), ("normal", g_norm, film.normal),
                       ("albedo", g_alb, film.albedo)):
        ok = relative_ok(g, c)
        results[name] = ok.mean()
        log(f"  {name}: fraction within {REL_TOL}: {ok.mean():.6f}; "
            f"bit-identical {np.mean(np.all(g.view(np.uint32) == c.view(np.uint32), axis=-1)):.6f}")
    return results


def test_film_parity_cornell_full(backend):
    """Cornell box at its full 640x400, 16 spp."""
    log("case: film parity cornell 640x400x16")
    r = film_parity(backend, scenes.cornell_box(), scenes.cornell_settings(), 0, 16, 16)
    for name, frac in r.items():
        assert frac >= PIXEL_FRACTION, name


def test_film_parity_instanced_small(backend):
    log("case: film parity small instanced 128x72x8")
    r = film_parity(backend, small_instanced(), scenes.instanced_settings(128, 72), 0, 8, 8)
    for name, frac in r.items():
        assert frac >= PIXEL_FRACTION, name


def test_film_parity_cornell_variants(backend):
    """Other integrator settings: forward-only, pure NEE, RR from the camera,
    clamping, transparent background, blackman-harris filter."""
    snap = scenes.cornell_box()
    variants = []
    s = scenes.cornell_settings(64, 40)
    s.integrator.direct_light_sampling = abi.DIRECT_LIGHT_FORWARD
    variants.append(("forward", s))
    s = scenes.cornell_settings(64, 40)
    s.integrator.direct_light_sampling = abi.DIRECT_LIGHT_NEE
    variants.append(("nee", s))
    s = scenes.cornell_settings(64, 40)
    s.integrator.min_bounces = 0
    s.integrator.sample_clamp_indirect = 0.5
    s.integrator.sample_clamp_direct = 4.0
    variants.append(("rr+clamp", s))
    s = scenes.cornell_settings(64, 40)
    s.transparent_background = 1
    s.pixel_filter = abi.FILTER_BLACKMAN_HARRIS
    s.filter_width = 1.5
    variants.append(("filter+alpha", s))
    for name, s in variants:
        log(f"case: cornell variant {name}")
        r = film_parity(backend, snap, s, 0, 8, 8)
        for k, frac in r.items():
            assert frac >= PIXEL_FRACTION, (name, k)


def test_film_parity_windowed_offset_range(backend):
    """A pixel window and a sample range that starts mid-sequence."""
    log("case: window + offset range")
    s = scenes.cornell_settings(96, 60)
    s.window_x, s.window_y, s.window_width, s.window_height = 16, 8, 40, 30
    r = film_parity(backend, scenes.cornell_box(), s, 5, 6, 32)
    for k, frac in r.items():
        assert frac >= PIXEL_FRACTION, k


# ------------------------------------------------------------------ contract

def test_progressive_chunks_bit_identical(backend):
    log("case: progressive chunks")
    snap = scenes.cornell_box()
    s = scenes.cornell_settings(64, 40)
    scene = compile_scene(backend, snap)
    a = backend.create_session(scene, s)
    a.render_device(0, 16, 16)
    b = backend.create_session(scene, s)
    b.render_device(0, 3, 16)
    b.render_device(3, 13, 16)
    fa, fb = a.read_film(), b.read_film()
    for x, y in zip(fa, fb):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_dispatch_batch_size_does_not_change_result(backend):
    log("case: batch size invariance")
    snap = scenes.cornell_box()
    s = scenes.cornell_settings(48, 30)
    films = []
    for per in (1, 4, 16):
        be = RendererBackend(0, max_samples_per_dispatch=per)
        scene = compile_scene(be, snap)
        sess = be.create_session(scene, s)
        sess.render_device(0, 16, 16)
        films.append(sess.read_film()[0])
    assert np.array_equal(films[0].view(np.uint32), films[1].view(np.uint32))
    assert np.array_equal(films[0].view(np.uint32), films[2].view(np.uint32))


def test_render_samples_sink_and_passes(backend):
    log("case: render_samples sink")
    snap = scenes.cornell_box()
    s, keep = with_passes(scenes.cornell_settings(64, 40))
    scene = compile_scene(backend, snap)
    session = backend.create_session(scene, s)
    sink = MemoryOutputSink()
    assert session.render_samples(SampleRange(0, 8, 0, 8), sink)
    assert sink.began and sink.ended is False
    assert set(sink.passes) == {abi.PASS_COMBINED, abi.PASS_NORMAL, abi.PASS_ALBEDO,
                                abi.PASS_SAMPLE_COUNT}
    assert sink.passes[abi.PASS_COMBINED].shape == (40, 64, 4)
    assert np.all(sink.passes[abi.PASS_SAMPLE_COUNT] == 8.0)
    film, _ = ob.OracleScene(snap).render(s, 0, 8, 8)
    expected = film.write_pass(abi.PASS_COMBINED, 4)
    assert relative_ok(sink.passes[abi.PASS_COMBINED], expected).mean() >= PIXEL_FRACTION
    # total must stay fixed for the session; ranges past it fail
    assert not session.render_samples(SampleRange(8, 4, 0, 16), MemoryOutputSink())
    assert session.last_status == abi.PSX_ERROR_RANGE
    assert not session.render_samples(SampleRange(0, 0, 0, 8), MemoryOutputSink())
    assert not session.render_samples(SampleRange(6, 4, 0, 8), MemoryOutputSink())
    # cancel sticks until reset
    session.cancel()
    assert not session.render_samples(SampleRange(0, 1, 0, 8), MemoryOutputSink())
    assert session.last_status == abi.PSX_ERROR_CANCELLED
    session.reset(s)
    assert session.render_samples(SampleRange(0, 2, 0, 8), MemoryOutputSink())


def test_invalid_session_settings(backend):
    snap = scenes.cornell_box()
    scene = compile_scene(backend, snap)
    s = scenes.cornell_settings(64, 40)
    s.width = 0
    assert backend.create_session(scene, s) is None
    s = scenes.cornell_settings(64, 40)
    s.window_x, s.window_width = 60, 10
    assert backend.create_session(scene, s) is None


def test_parameter_update_matches_recompile(backend):
    log("case: parameter-only rebind")
    snap = scenes.cornell_box()
    s = scenes.cornell_settings(64, 40)
    scene = compile_scene(backend, snap)
    new = snap.materials[2].parameters.copy()
    new[:3] = (0.9, 0.2, 0.7)
    assert scene.update_parameters(2, new)
    assert not scene.update_parameters(2, new[:2])
    a = backend.create_session(scene, s)
    a.render_device(0, 4, 4)
    snap2 = scenes.cornell_box()
    snap2.materials[2].parameters = new
    b = backend.create_session(compile_scene(backend, snap2), s)
    b.render_device(0, 4, 4)
    assert np.array_equal(a.read_film()[0].view(np.uint32), b.read_film()[0].view(np.uint32))


def test_compile_diagnostics(backend):
    snap = scenes.cornell_box()
    snap.instances[0].mesh = 99
    comp = backend.compile_scene(snap)
    assert not comp.ok()
    assert comp.diagnostics and "mesh" in comp.diagnostics[0].message


# ------------------------------------------------------------------ full size

def test_instanced_full_scene_rows_match_oracle(backend):
    """The 1,003,520-triangle icosphere instanced 64 times at 1920x1080:
    a band of rows at 2 spp agrees with the oracle's own BVH render."""
    log("case: full instanced scene, rows 500-507 at 2 spp")
    snap = scenes.instanced_icospheres()
    s = scenes.instanced_settings()
    scene = compile_scene(backend, snap)
    session = backend.create_session(scene, s)
    session.render_device(0, 2, 64)
    g_comb, g_norm, g_alb, g_cnt = session.read_film()
    assert np.all(g_cnt == 2)
    assert np.all(np.isfinite(g_comb))
    film = ob.Film(1920, 1080)
    ob.OracleScene(snap).render(s, 0, 2, 64, film=film, rows=(500, 508))
    rows = slice(500, 508)
    for name, g, c in (("combined", g_comb, film.combined), ("albedo", g_alb, film.albedo),
                       ("normal", g_norm, film.normal)):
        ok = relative_ok(g[rows], c[rows])
        log(f"  {name}: {ok.mean():.6f}")
        assert ok.mean() >= PIXEL_FRACTION
