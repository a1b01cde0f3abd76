"""BVH branching factor (PSX_BACKEND_BVH_WIDTH_*) on the GPU.

Closest hits resolve ties by (object, primitive), not by the tree, so films
and path events are the same whatever the width: every film at widths 2 and
3 must be bit-identical to width 4, for both integrators.  The scene-level
tuner (psycles_amd.tuning.tune_bvh_width) times each width and keeps four
unless another is at least 1 % faster.
"""
import sys

import numpy as np
import pytest

from psycles_amd import abi, scenes, tuning
from psycles_amd.render import RendererBackend

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def film(snap, s, width, integrator, spp=4):
    be = RendererBackend(0, integrator=integrator, bvh_width=width)
    comp = be.compile_scene(snap)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    sess = be.create_session(comp.scene, s)
    sess.render_device(0, spp, spp)
    out = sess.read_film()
    st = sess.stats()
    sess.close()
    comp.scene.close()
    be.close()
    return out, st.rays


@pytest.mark.parametrize("case", ["cornell_lamps", "instanced", "many_lights"])
@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_WAVEFRONT, abi.INTEGRATOR_MEGAKERNEL])
def test_films_identical_across_widths(case, integrator):
    if case == "cornell_lamps":
        snap, s = scenes.cornell_lamps(), scenes.cornell_settings(48, 30)
    elif case == "instanced":
        snap, s = scenes.instanced_icospheres(frequency=12, instances=8), scenes.instanced_settings(64, 36)
    else:
        snap = scenes.many_lights(frequency=6, instances=6, emissive=150, points=12, spots=4)
        s = scenes.many_light_settings(64, 36)
    ref, rays = film(snap, s, 4, integrator)
    for width in (2, 3):
        got, got_rays = film(snap, s, width, integrator)
        log(f"case: {case}, integrator {integrator}, width {width}: rays {got_rays} vs {rays}")
        assert got_rays == rays
        for x, y in zip(got, ref):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (case, width)


def test_both_width_flags_are_rejected():
    import ctypes as C
    lib = abi.load_library()
    opts = abi.BackendOptions(1, 4, abi.INTEGRATOR_WAVEFRONT, abi.BACKEND_BVH_WIDTH_2 | abi.BACKEND_BVH_WIDTH_3)
    handle = C.c_void_p()
    assert lib.psx_backend_create(0, C.byref(opts), C.byref(handle)) == abi.PSX_ERROR_INVALID_ARGUMENT
    assert not handle.value


def test_tuner_picks_a_width():
    snap, s = scenes.instanced_icospheres(frequency=12, instances=8), scenes.instanced_settings(128, 72)
    r = tuning.tune_bvh_width(snap, s, probe_samples=2)
    log(f"case: BVH width tuner: {r}")
    assert r["chosen"] in (2, 3, 4)
    assert all(ms > 0.0 for ms in r["ms"].values())
    assert r["ms"][r["chosen"]] <= r["ms"][4]
