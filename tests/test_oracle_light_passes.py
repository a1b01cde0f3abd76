"""Light passes in the CPU oracle (src/luisa/path_kernel_setup.cpp:696-733,
src/luisa/path_tracer_lighting.cpp:107-191, src/luisa/path_tracer_session.cpp:150-230).

Pinned by what the pass definitions fix: every clamped contribution lands in
exactly one of emission, environment or the diffuse / glossy / transmission
direct-indirect split, so their sums equal Combined; closures evaluated here
are diffuse, emission and transparent, so glossy, transmission, volume and
colour passes are zero; a path that may not bounce has no indirect light;
a directly seen emitter's emission pass is its strength; the divided
diffuse passes times the albedo give back the sums.
"""
import numpy as np

import oracle_binding as ob
from psycles_amd import abi, scenes

DD, DI, GD, GI, TD, TI, VD, VI, EM, ENV, GC, TC = range(12)


def render(snap, s, spp=4):
    film, _ = ob.OracleScene(snap).render(s, 0, spp, spp, light=True)
    return film


def test_passes_sum_to_combined_and_zero_lobes():
    snap = scenes.cornell_lamps()
    s = scenes.cornell_settings(48, 32)
    film = render(snap, s)
    lp = film.light
    total = lp[..., [DD, DI, GD, GI, TD, TI, EM, ENV], :3].sum(axis=2)
    comb = film.combined[..., :3]
    assert np.allclose(total, comb, rtol=1e-5, atol=1e-6)
    for k in (GD, GI, VD, VI, GC, TC):
        assert np.all(lp[..., k, :3] == 0.0)
    assert np.all(np.abs(lp[..., [TD, TI], :3]) <= 1e-6 * (np.abs(comb) + 1e-6)[..., None, :])
    assert lp[..., DD, :3].sum() > 0 and lp[..., DI, :3].sum() > 0
    assert np.all(lp[..., :, 3] == 4.0)


def test_transparent_materials_and_background_route():
    snap = scenes.instanced_icospheres(frequency=8, instances=6)
    s = scenes.instanced_settings(64, 36)
    film = render(snap, s)
    lp = film.light
    total = lp[..., [DD, DI, GD, GI, TD, TI, EM, ENV], :3].sum(axis=2)
    assert np.allclose(total, film.combined[..., :3], rtol=1e-5, atol=1e-6)
    # the constant world is seen directly (through transparency too) and after bounces
    assert lp[..., ENV, :3].sum() > 0 and lp[..., DI, :3].sum() > 0


def test_no_bounce_has_no_indirect_light():
    snap = scenes.cornell_box()
    s = scenes.cornell_settings(48, 32)
    s.integrator.max_bounces = 0
    film = render(snap, s)
    lp = film.light
    assert np.all(lp[..., DI, :3] == 0.0)
    assert lp[..., DD, :3].sum() > 0


def test_directly_seen_emitter_and_divided_passes():
    snap = scenes.cornell_box()
    s = scenes.cornell_settings(160, 100)
    s.integrator.direct_light_sampling = abi.DIRECT_LIGHT_FORWARD
    film = render(snap, s)
    emission = film.write_light_pass(abi.PASS_EMISSION, 3)
    lit = emission.max(axis=-1) > 0
    assert lit.any()
    # the ceiling quad's radiance is its strength, 15: each of the 4 samples
    # of a pixel sees 15 or nothing
    quarters = emission[lit] / 15.0 * 4.0
    assert np.allclose(quarters, np.round(quarters), atol=1e-4)
    assert np.isclose(emission[lit].max(), 15.0, rtol=1e-6)
    dd = film.write_light_pass(abi.PASS_DIFFUSE_DIRECT, 3)
    alb = film.albedo[..., :3]
    full = np.all(alb > 0, axis=-1)
    assert np.allclose((dd * alb)[full], film.light[..., DD, :3][full], rtol=1e-5, atol=1e-6)
