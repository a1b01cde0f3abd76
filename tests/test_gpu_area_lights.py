"""Area lamps and the orthographic camera on the GPU, through the C ABI.

* The Cycles renders held by tests/test_luisa_area_light_forward.cpp (a
  rectangle lamp with full spread, flat CDF and light tree, and a
  narrow-spread ellipse, seen by an orthographic camera): frame means and
  recorded pixels within that test's tolerances, on both integrators.
* The Cornell box lit by an analytic 0.25 x 0.25 area lamp in place of the
  emissive quad (the other reading of the Cornell configuration), a
  narrow-spread rectangle and an ellipse: Combined / Normal / Albedo within
  1e-4 relative of the oracle on >= 99.9 % of pixels, discrete path events
  (flags, RNG offset, hit identity, light and closure choice, termination
  bits) identical on every event, in MIS, NEE-only and forward-only
  modes, flat CDF and light tree.
* Orthographic primary hits identical to the oracle's.
"""
import math
import sys

import numpy as np
import pytest

import event_flips
import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.render import RendererBackend
from test_oracle_area_lights import ELLIPSE_MEANS, ELLIPSE_PIXELS, RECT_MEANS, RECT_PIXELS, fixture

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4
PIXEL_FRACTION = 0.999


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relative_ok(a, b, tol=REL_TOL):
    diff = np.abs(a - b).max(axis=-1)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)).max(axis=-1), 1e-3)
    return diff <= tol * scale


INTEGRATORS = [abi.INTEGRATOR_MEGAKERNEL, abi.INTEGRATOR_WAVEFRONT]


@pytest.mark.parametrize("integrator", INTEGRATORS, ids=["megakernel", "wavefront"])
@pytest.mark.parametrize("case", ["rectangle", "rectangle_tree", "narrow_ellipse"])
def test_cycles_area_light_fixture(case, integrator):
    log(f"case: Cycles area-light fixture {case}, integrator {integrator}")
    ellipse = case == "narrow_ellipse"
    snap, s = fixture(ellipse, case == "rectangle_tree")
    be = RendererBackend(0, integrator=integrator)
    comp = be.compile_scene(snap)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    sess = be.create_session(comp.scene, s)
    sess.render_device(0, 1, 1)
    c = sess.read_film()[0][..., :3].astype(np.float64)
    means, pixels = (ELLIPSE_MEANS, ELLIPSE_PIXELS) if ellipse else (RECT_MEANS, RECT_PIXELS)
    m = c.reshape(-1, 3).mean(axis=0)
    log(f"  means {m} (Cycles {means})")
    assert np.all(np.abs(m - np.asarray(means)) <= 2.0e-6)
    for (x, y), v in pixels.items():
        for a, e in zip(c[y, x], v):
            assert abs(a - e) <= 2.0e-6 * (1.0 + max(abs(a), abs(e))), ((x, y), c[y, x], v)


def cornell_area(kind):
    """The Cornell box with the ceiling quad replaced by an analytic area
    lamp of the same size and radiance (strength 15 over 0.25 x 0.25 at
    y = 0.999 facing -y), or a narrow-spread / elliptic variant."""
    snap = scenes.cornell_box()
    snap.instances = snap.instances[:3]
    # lamp z axis +y, so its emitting side (-z) faces the floor
    frame = scenes.frame_z((0.0, 1.0, 0.0))
    side = 0.25
    power = 15.0 * side * side * math.pi   # normalised power: radiance 15
    l = scenes.lamp(abi.LIGHT_AREA, scenes.translation(0.0, 0.999, 0.0) @ frame, (1.0, 1.0, 1.0), power,
                    size=side, sphere=False)
    l.size_y = side
    if kind == "narrow":
        l.spread = 0.7
        l.size_y = 0.4
    elif kind == "ellipse":
        l.flags |= abi.LIGHT_FLAG_ELLIPSE
        l.spread = 1.1
        l.size = 0.3
        l.size_y = 0.3
    snap.lights = [l]
    snap.name = f"cornell_area_{kind}"
    return snap


@pytest.mark.parametrize("integrator", INTEGRATORS, ids=["megakernel", "wavefront"])
@pytest.mark.parametrize("mode", [abi.DIRECT_LIGHT_MIS, abi.DIRECT_LIGHT_NEE, abi.DIRECT_LIGHT_FORWARD],
                         ids=["mis", "nee", "forward"])
@pytest.mark.parametrize("kind", ["square", "narrow", "ellipse"])
def test_cornell_area_lamp_film(kind, mode, integrator):
    log(f"case: cornell area lamp {kind}, mode {mode}, integrator {integrator}")
    snap = cornell_area(kind)
    for tree in (0, 1):
        s = scenes.cornell_settings(64, 40)
        s.integrator.direct_light_sampling = mode
        s.integrator.use_light_tree = tree
        be = RendererBackend(0, integrator=integrator)
        comp = be.compile_scene(snap)
        assert comp.ok(), [d.message for d in comp.diagnostics]
        sess = be.create_session(comp.scene, s)
        sess.render_device(0, 8, 8)
        g = sess.read_film()
        film, _ = ob.OracleScene(snap).render(s, 0, 8, 8)
        assert np.array_equal(g[3], film.sample_count)
        for name, a, b in (("combined", g[0], film.combined), ("normal", g[1], film.normal),
                           ("albedo", g[2], film.albedo)):
            frac = relative_ok(a, b).mean()
            log(f"  tree {tree} {name}: {frac:.6f}")
            assert frac >= PIXEL_FRACTION, (tree, name)


def test_cornell_area_lamp_events():
    log("case: cornell area lamp path events")
    snap = cornell_area("narrow")
    s = scenes.cornell_settings(48, 30)
    be = RendererBackend(0)
    sess = be.create_session(be.compile_scene(snap).scene, s)
    osc = ob.OracleScene(snap)
    event_flips.check("cornell area lamp", sess, osc, s, (0, 5), 8, log)


def test_orthographic_primary_hits():
    log("case: orthographic primary hits")
    snap = scenes.instanced_icospheres(frequency=12, instances=8)
    snap.camera = scenes.CameraData(scenes.look_at((0.0, 6.0, 58.0), (0.0, -2.0, 0.0)), math.radians(42.0),
                                    1.0e-3, 1000.0, sensor_fit=2, projection=1, orthographic_scale=60.0)
    s = scenes.instanced_settings(96, 54)
    be = RendererBackend(0)
    sess = be.create_session(be.compile_scene(snap).scene, s)
    osc = ob.OracleScene(snap)
    for sample in (0, 3):
        gpu = sess.primary_hits(sample, 8)
        cpu = osc.primary_hits(s, sample, 8)
        assert np.array_equal(gpu, cpu)
        assert (gpu[..., 3] != 0).mean() > 0.05
