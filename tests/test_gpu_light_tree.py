"""Light-tree NEE on the GPU vs the CPU oracle, through the C ABI.

The tree (src/sampling/light_tree.cpp build, src/luisa/path_tracer_light_tree.cpp
descent and reverse pdf) picks the light for NEE and supplies the selection
pdf of forward MIS for emissive triangles and lamps.  Scenes: the Cornell box
with lamps, and a small many-light scene (instanced icospheres, emissive
add(emission, diffuse) triangles, point / spot / distant lamps).  Bars as for
the rest of the path: light choice and other discrete events identical,
Combined / Normal / Albedo within 1e-4 relative on >= 99.9 % of pixels.
"""
import sys

import numpy as np
import pytest

import event_flips
import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.render import RendererBackend

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4
PIXEL_FRACTION = 0.999
INTEGRATORS = [abi.INTEGRATOR_MEGAKERNEL, abi.INTEGRATOR_WAVEFRONT]
INTEGRATOR_IDS = ["megakernel", "wavefront"]
DISCRETE = [abi.TRACE_FLAGS, abi.TRACE_RNG_OFFSET, abi.TRACE_OBJECT, abi.TRACE_PRIMITIVE,
            abi.TRACE_LIGHT, abi.TRACE_CLOSURE, abi.TRACE_TERMINATE_BITS]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relative_ok(a, b, tol=REL_TOL):
    diff = np.abs(a - b).max(axis=-1)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)).max(axis=-1), 1e-3)
    return diff <= tol * scale


def film_vs_oracle(backend, snap, s, first, count, total):
    comp = backend.compile_scene(snap)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    session = backend.create_session(comp.scene, s)
    session.render_device(first, count, total)
    g = session.read_film()
    film, _ = ob.OracleScene(snap).render(s, first, count, total)
    assert np.array_equal(g[3], film.sample_count)
    for name, a, b in (("combined", g[0], film.combined), ("normal", g[1], film.normal),
                       ("albedo", g[2], film.albedo)):
        frac = relative_ok(a, b).mean()
        log(f"  {name}: {frac:.6f}")
        assert frac >= PIXEL_FRACTION, name
    return g


def small_many_lights():
    return scenes.many_lights(frequency=6, instances=6, emissive=150, points=12, spots=4)


def tree_settings(w, h, mode=abi.DIRECT_LIGHT_MIS, cornell=True):
    s = scenes.cornell_settings(w, h) if cornell else scenes.many_light_settings(w, h)
    s.integrator.direct_light_sampling = mode
    s.integrator.use_light_tree = 1
    return s


@pytest.mark.parametrize("integrator", INTEGRATORS, ids=INTEGRATOR_IDS)
@pytest.mark.parametrize("mode", [abi.DIRECT_LIGHT_MIS, abi.DIRECT_LIGHT_NEE, abi.DIRECT_LIGHT_FORWARD],
                         ids=["mis", "nee", "forward"])
def test_cornell_lamps_tree_film(integrator, mode):
    log(f"case: cornell lamps, light tree, integrator {integrator}, mode {mode}")
    be = RendererBackend(0, integrator=integrator)
    film_vs_oracle(be, scenes.cornell_lamps(), tree_settings(64, 40, mode), 0, 8, 8)


def test_cornell_mesh_light_only_tree_film():
    log("case: cornell mesh light, light tree")
    be = RendererBackend(0)
    film_vs_oracle(be, scenes.cornell_box(), tree_settings(64, 40), 0, 8, 8)


@pytest.mark.parametrize("integrator", INTEGRATORS, ids=INTEGRATOR_IDS)
def test_many_lights_tree_film(integrator):
    log(f"case: small many-light scene, light tree, integrator {integrator}")
    be = RendererBackend(0, integrator=integrator)
    film_vs_oracle(be, small_many_lights(), tree_settings(64, 40, cornell=False), 0, 4, 4)


def test_many_lights_flat_film():
    log("case: small many-light scene, flat distribution")
    be = RendererBackend(0)
    s = tree_settings(64, 40, cornell=False)
    s.integrator.use_light_tree = 0
    film_vs_oracle(be, small_many_lights(), s, 0, 4, 4)


@pytest.mark.parametrize("snap_name", ["cornell_lamps", "many_lights"])
def test_tree_path_events(snap_name):
    log(f"case: light tree path events, {snap_name}")
    if snap_name == "cornell_lamps":
        snap, s = scenes.cornell_lamps(), tree_settings(80, 50)
    else:
        snap, s = small_many_lights(), tree_settings(80, 50, cornell=False)
    be = RendererBackend(0)
    comp = be.compile_scene(snap)
    session = be.create_session(comp.scene, s)
    osc = ob.OracleScene(snap)
    event_flips.check(f"light tree {snap_name}", session, osc, s, (0, 3), 6, log)


def test_tree_integrators_bit_identical():
    log("case: light tree, wavefront vs megakernel")
    snap, s = small_many_lights(), tree_settings(96, 60, cornell=False)
    films = []
    for integ in INTEGRATORS:
        be = RendererBackend(0, integrator=integ)
        comp = be.compile_scene(snap)
        sess = be.create_session(comp.scene, s)
        sess.render_device(0, 4, 4)
        films.append((sess.read_film(), sess.stats().rays))
    for a, b in zip(films[0][0], films[1][0]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert films[0][1] == films[1][1]


def test_tree_parameter_edit_rebuilds_tree():
    """A strength edit changes emitter energies; the tree is rebuilt from the
    new parameters and the film still matches the oracle of the edited scene."""
    log("case: light tree after a parameter-only strength edit")
    be = RendererBackend(0)
    snap = scenes.cornell_lamps()
    s = tree_settings(64, 40)
    comp = be.compile_scene(snap)
    assert comp.ok()
    session = be.create_session(comp.scene, s)
    light_material = next(i for i, m in enumerate(snap.materials) if "light" in m.name.lower())
    params = snap.materials[light_material].parameters.copy()
    params[-1] *= 0.25
    snap.materials[light_material].parameters = params
    comp.scene.update_parameters(light_material, params)
    session.render_device(0, 8, 8)
    g = session.read_film()
    film, _ = ob.OracleScene(snap).render(s, 0, 8, 8)
    frac = relative_ok(g[0], film.combined).mean()
    log(f"  combined: {frac:.6f}")
    assert frac >= PIXEL_FRACTION


def test_many_lights_full_scene_rows_match_oracle():
    """The full many-light scene (64 instanced 1,003,520-triangle icospheres,
    10,000 emissive triangles, 321 lamps, light tree) at 1920x1080: a band of
    rows at 1 spp of 32 agrees with the oracle's own BVH render."""
    log("case: full many-light scene, rows 600-603 at 1 spp")
    snap = scenes.many_lights()
    s = scenes.many_light_settings()
    be = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)
    comp = be.compile_scene(snap)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    session = be.create_session(comp.scene, s)
    session.render_device(0, 1, 32)
    g_comb, g_norm, g_alb, g_cnt = session.read_film()
    assert np.all(g_cnt == 1)
    assert np.all(np.isfinite(g_comb))
    film = ob.Film(1920, 1080)
    ob.OracleScene(snap).render(s, 0, 1, 32, film=film, rows=(600, 604))
    rows = slice(600, 604)
    for name, g, c in (("combined", g_comb, film.combined), ("albedo", g_alb, film.albedo),
                       ("normal", g_norm, film.normal)):
        ok = relative_ok(g[rows], c[rows])
        log(f"  {name}: {ok.mean():.6f}")
        assert ok.mean() >= PIXEL_FRACTION
