"""Scenes loaded from Blender bundles, rendered on the GPU vs the oracle.

The bundle is written from a SceneSnapshot (v1 point normals and v2 corner
normals), loaded with psycles_amd.blender and compiled through the C ABI;
bars as for the rest of the path: Combined / Normal / Albedo within 1e-4
relative on >= 99.9 % of pixels, sample counts identical.
"""
import sys

import numpy as np
import pytest

import blender_bundle as bb
import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.blender import load_blender_scene_bundle
from psycles_amd.render import RendererBackend

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


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_MEGAKERNEL, abi.INTEGRATOR_WAVEFRONT],
                         ids=["megakernel", "wavefront"])
@pytest.mark.parametrize("scene", ["smooth_mixed_v2", "cornell_lamps_v1"])
def test_bundle_gpu_vs_oracle(tmp_path, integrator, scene):
    log(f"bundle {scene} integrator {integrator}")
    if scene == "smooth_mixed_v2":
        snap, s, compact = bb.smooth_sphere_snapshot(), abi.default_settings(96, 64), True
        s.seed = 11
    else:
        snap, s, compact = scenes.cornell_lamps(), scenes.cornell_settings(96, 60), False
    bb.snapshot_bundle(tmp_path / "b", snap, s, samples=8, compact=compact)
    imported = load_blender_scene_bundle(tmp_path / "b")
    assert imported.ok(), [d.message for d in imported.diagnostics]
    loaded, settings = imported.snapshot, imported.settings
    backend = RendererBackend(integrator=integrator)
    comp = backend.compile_scene(loaded)
    assert comp.ok(), [d.message for d in comp.diagnostics]
    session = backend.create_session(comp.scene, settings)
    session.render_device(0, imported.samples, imported.samples)
    g = session.read_film()
    film, _ = ob.OracleScene(loaded).render(settings, 0, imported.samples, imported.samples)
    assert np.array_equal(g[3], film.sample_count)
    for name, a, b in (("combined", g[0], film.combined), ("normal", g[1], film.normal),
                       ("albedo", g[2], film.albedo)):
        frac = relative_ok(a, b).mean()
        log(f"  {name}: {frac:.6f}")
        assert frac >= PIXEL_FRACTION, name
    session.close()
    comp.scene.close()
    backend.close()
