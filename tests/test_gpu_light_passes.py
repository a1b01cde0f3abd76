"""Light passes on the GPU vs the CPU oracle, through the C ABI.

The twelve light-pass sums (diffuse / glossy / transmission / volume direct
and indirect, emission, environment, glossy and transmission colour) of
sessions that request light passes, compared with the oracle's sums pass by
pass within 1e-4 relative on >= 99.9 % of pixels; the Combined / Normal /
Albedo film is unchanged, bit for bit, by requesting them; the passes a sink
receives match the oracle's pass output.
"""
import sys

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.render import MemoryOutputSink, RendererBackend, SampleRange, with_passes
from psycles_amd.server import RenderServer

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4
PIXEL_FRACTION = 0.999
LIGHT = ((abi.PASS_COMBINED, 4), (abi.PASS_DIFFUSE_DIRECT, 3), (abi.PASS_DIFFUSE_INDIRECT, 3),
         (abi.PASS_GLOSSY_DIRECT, 3), (abi.PASS_TRANSMISSION_INDIRECT, 3), (abi.PASS_EMISSION, 3),
         (abi.PASS_ENVIRONMENT, 3), (abi.PASS_GLOSSY_COLOR, 3), (abi.PASS_ALBEDO, 3), (abi.PASS_SAMPLE_COUNT, 1))


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relative_ok(a, b, tol=REL_TOL):
    diff = np.abs(a - b).max(axis=-1)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)).max(axis=-1), 1e-3)
    return diff <= tol * scale


CASES = {
    "cornell_lamps_mis": (scenes.cornell_lamps, lambda: scenes.cornell_settings(64, 40), abi.DIRECT_LIGHT_MIS),
    "cornell_nee": (scenes.cornell_box, lambda: scenes.cornell_settings(64, 40), abi.DIRECT_LIGHT_NEE),
    "cornell_forward": (scenes.cornell_box, lambda: scenes.cornell_settings(64, 40), abi.DIRECT_LIGHT_FORWARD),
    "instanced": (lambda: scenes.instanced_icospheres(frequency=10, instances=8),
                  lambda: scenes.instanced_settings(96, 54), abi.DIRECT_LIGHT_MIS),
}


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_MEGAKERNEL, abi.INTEGRATOR_WAVEFRONT],
                         ids=["megakernel", "wavefront"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_light_pass_sums_match_oracle(case, integrator):
    log(f"case: light passes, {case}, integrator {integrator}")
    make_scene, make_settings, mode = CASES[case]
    snap = make_scene()
    s, keep = with_passes(make_settings(), LIGHT)
    s.integrator.direct_light_sampling = mode
    be = RendererBackend(0, integrator=integrator)
    comp = be.compile_scene(snap)
    sess = be.create_session(comp.scene, s)
    sess.render_device(0, 8, 8)
    gpu = sess.read_light_passes()
    g_comb, g_norm, g_alb, g_cnt = sess.read_film()
    film, _ = ob.OracleScene(snap).render(s, 0, 8, 8, light=True)
    assert np.array_equal(gpu[..., 3], film.light[..., 3])
    for b in range(12):
        frac = relative_ok(gpu[..., b, :3], film.light[..., b, :3]).mean()
        if frac < 1.0:
            log(f"  pass slot {b}: {frac:.6f}")
        assert frac >= PIXEL_FRACTION, b
    # requesting light passes leaves the film as it is
    s2 = make_settings()
    s2.integrator.direct_light_sampling = mode
    p_sess = be.create_session(comp.scene, s2)
    p_sess.render_device(0, 8, 8)
    for a, b in zip((g_comb, g_norm, g_alb, g_cnt), p_sess.read_film()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_sink_light_passes_match_oracle():
    log("case: light passes through a sink")
    snap = scenes.cornell_lamps()
    s, keep = with_passes(scenes.cornell_settings(64, 40), LIGHT)
    be = RendererBackend(0)
    comp = be.compile_scene(snap)
    sess = be.create_session(comp.scene, s)
    sink = MemoryOutputSink()
    assert sess.render_samples(SampleRange(0, 8, 0, 8), sink)
    film, _ = ob.OracleScene(snap).render(s, 0, 8, 8, light=True)
    for kind, channels in LIGHT:
        if kind in (abi.PASS_COMBINED, abi.PASS_ALBEDO, abi.PASS_SAMPLE_COUNT):
            continue
        expect = film.write_light_pass(kind, channels)
        frac = relative_ok(sink.passes[kind], expect).mean()
        log(f"  pass {kind}: {frac:.6f}")
        assert frac >= PIXEL_FRACTION, kind


def test_server_rejects_light_passes():
    be = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)
    comp = be.compile_scene(scenes.cornell_box())
    s, keep = with_passes(scenes.cornell_settings(32, 32), LIGHT)
    server = RenderServer(be, path_capacity=1 << 14, streams=2)
    with pytest.raises(abi.PsxError):
        server.submit(comp.scene, s, 0, 1, 1)
    server.close()
