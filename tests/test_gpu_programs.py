"""Generated surface programs on the device: wavefront films shaded with
the scene's hipRTC-compiled programs are bit-identical to the interpreter's,
for every shade variant, across parameter rebinds (which reuse the compiled
module) and under graph replay."""
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
def generated():
    return RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=4,
                           generated_programs=True)


@pytest.fixture(scope="module")
def interpreter():
    return RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=4)


def oren_nayar_cornell():
    snap = scenes.cornell_box()
    for i, sigma in ((0, 0.5), (1, 0.8), (2, 0.3)):
        m = snap.materials[i]
        snap.materials[i] = scenes.diffuse_material(tuple(m.parameters[:3]), roughness=sigma)
    return snap


CASES = {
    "cornell": (scenes.cornell_box, lambda: scenes.cornell_settings(64, 40)),
    "cornell_lamps": (scenes.cornell_lamps, lambda: scenes.cornell_settings(64, 40)),
    "oren_nayar": (oren_nayar_cornell, lambda: scenes.cornell_settings(64, 40)),
    "instanced": (lambda: scenes.instanced_icospheres(frequency=10, instances=12, box=14.0),
                  lambda: scenes.instanced_settings(96, 54)),
    "many_lights": (lambda: scenes.many_lights(frequency=6, instances=6, emissive=150, points=12, spots=4),
                    lambda: scenes.many_light_settings(96, 54)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_generated_programs_match_the_interpreter(generated, interpreter, name):
    make_scene, make_settings = CASES[name]
    snap, s = make_scene(), make_settings()
    log(f"case: generated surface programs vs interpreter, {name}")
    before = generated.generated_programs
    ga = generated.create_session(generated.compile_scene(snap).scene, s)
    ia = interpreter.create_session(interpreter.compile_scene(snap).scene, s)
    for first, count in ((0, 4), (4, 3)):
        ga.render_device(first, count, 8)
        ia.render_device(first, count, 8)
        assert_bit_identical(ga.read_film(), ia.read_film(), (name, first))
    log(f"  identical; modules compiled {generated.generated_programs - before}")
    assert generated.generated_programs >= max(before, 1)   # the generated kernels ran
    assert interpreter.generated_programs == 0


def test_parameter_rebind_reuses_the_module(generated, interpreter):
    snap, s = scenes.cornell_box(), scenes.cornell_settings(48, 30)
    gs = generated.compile_scene(snap).scene
    isc = interpreter.compile_scene(snap).scene
    a = generated.create_session(gs, s)
    a.render_device(0, 4, 4)
    compiled = generated.generated_programs
    assert compiled >= 1
    new = snap.materials[2].parameters.copy()
    new[:3] = (0.9, 0.2, 0.7)
    assert gs.update_parameters(2, new) and isc.update_parameters(2, new)
    a.reset(s)
    b = interpreter.create_session(isc, s)
    a.render_device(0, 4, 4)
    b.render_device(0, 4, 4)
    assert generated.generated_programs == compiled
    assert_bit_identical(a.read_film(), b.read_film(), "rebind")


def test_generated_programs_with_light_passes(generated, interpreter):
    light = ((abi.PASS_COMBINED, 4), (abi.PASS_DIFFUSE_DIRECT, 3), (abi.PASS_DIFFUSE_INDIRECT, 3),
             (abi.PASS_EMISSION, 3), (abi.PASS_ALBEDO, 3), (abi.PASS_SAMPLE_COUNT, 1))
    s, keep = with_passes(scenes.cornell_settings(40, 24), light)
    snap = scenes.cornell_lamps()
    log("case: generated programs, light-pass variant")
    ga = generated.create_session(generated.compile_scene(snap).scene, s)
    ia = interpreter.create_session(interpreter.compile_scene(snap).scene, s)
    ga.render_device(0, 4, 4)
    ia.render_device(0, 4, 4)
    assert_bit_identical(ga.read_film(), ia.read_film(), "light passes")
    assert np.array_equal(ga.read_light_passes().view(np.uint32), ia.read_light_passes().view(np.uint32))
    assert keep is not None


def test_generated_programs_without_graphs():
    eager = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=4,
                            generated_programs=True, graphs=False, time_trace=True)
    plain = RendererBackend(0, integrator=abi.INTEGRATOR_MEGAKERNEL, max_samples_per_dispatch=4)
    snap, s = scenes.cornell_box(), scenes.cornell_settings(48, 30)
    a = eager.create_session(eager.compile_scene(snap).scene, s)
    b = plain.create_session(plain.compile_scene(snap).scene, s)
    a.render_device(0, 8, 8)
    b.render_device(0, 8, 8)
    assert a.stats().trace_launches > 0
    assert_bit_identical(a.read_film(), b.read_film(), "host launches vs megakernel")


def test_server_batches_use_generated_programs(generated, interpreter):
    from psycles_amd.server import RenderServer
    snap = scenes.cornell_lamps()
    gs = generated.compile_scene(snap).scene
    isc = interpreter.compile_scene(snap).scene
    server = RenderServer(generated)
    handles = []
    for seed, (w, h) in enumerate(((48, 30), (64, 40), (32, 32))):
        s = scenes.cornell_settings(w, h)
        s.seed = seed
        handles.append((s, server.submit(gs, s, 0, 4, 4)))
    for s, hd in handles:
        assert hd.wait(-1.0)
        got = hd.read_film()
        hd.release()
        ref = interpreter.create_session(isc, s)
        ref.render_device(0, 4, 4)
        assert_bit_identical(got, ref.read_film(), ("server", s.width))
    server.close()
