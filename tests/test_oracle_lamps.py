"""Analytic lamps in the CPU oracle: point, spot and distant lamps.

Pinned by the Cycles values the reference's own test holds
(tests/test_luisa_analytic_light_sampling.cpp:350-991), by the forward/NEE
consistency that test also checks, and by closed forms: a delta point lamp
and a zero-angle sun over a diffuse plane give exact direct illumination,
and a delta spot leaves pixels outside its cone black.
"""
import math

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, scenes


def close(actual, expected, rel):
    return abs(actual - expected) <= 1.0e-7 + rel * max(abs(expected), abs(actual))


def finite_lamp(case, g):
    kind = abi.LIGHT_SPOT if case["type"] == "spot" else abi.LIGHT_POINT
    return scenes.lamp(kind, scenes.translation(*g["center"]), size=case["radius"],
                       spot_angle=g["spot_angle"], spot_smooth=g["spot_smooth"],
                       normalize=g["normalize"], sphere=case["sphere"])


def test_finite_lamp_samples_match_cycles(golden):
    g = golden["analytic_lamps"]["finite"]
    rel = golden["analytic_lamps"]["tolerance_relative"]
    for case in g["cases"]:
        light = finite_lamp(case, g)
        s = ob.sample_lamp(light, g["reference"], g["shading_normal"], False, g["random"])
        assert s["valid"], case
        got = {"direction_pdf": s["direction"] + [s["pdf"]],
               "position_distance": s["position"] + [s["distance"]],
               "normal_eval": s["normal"] + [s["eval"]]}
        for key, values in got.items():
            for a, e in zip(values, case[key]):
                assert close(a, e, rel), (case["type"], case["sphere"], key, values, case[key])


def test_forward_hit_agrees_with_sample(golden):
    """A forward hit along the sampled direction reproduces its distance and
    directional pdf (same check as the reference test)."""
    g = golden["analytic_lamps"]["finite"]
    for case in g["cases"]:
        light = finite_lamp(case, g)
        s = ob.sample_lamp(light, g["reference"], g["shading_normal"], False, g["random"])
        h = ob.intersect_lamp(light, g["reference"], s["direction"], 0.0, 1.0e10,
                              g["shading_normal"], False)
        assert h["valid"]
        assert close(h["distance"], s["distance"], 1.0e-5)
        assert close(h["pdf"], s["pdf"], 2.0e-6)
        assert close(h["eval"], s["eval"], 2.0e-6)


def test_distant_lamp_matches_cycles(golden):
    g = golden["analytic_lamps"]["distant"]
    axis = np.asarray(g["axis_z"])
    light = scenes.lamp(abi.LIGHT_DISTANT, scenes.frame_z(axis), angle=g["angle"])
    # frame_z normalises in float64: feed the exact float32 axis instead
    for r in range(3):
        light.transform[8 + r] = float(np.float32(axis[r]))
    s = ob.sample_lamp(light, [0, 0, 0], [0, 0, 1], False, g["random"])
    for a, e in zip(s["direction"] + [s["pdf"]], g["direction_pdf"]):
        assert close(a, e, 2.0e-6), (s, g["direction_pdf"])
    assert close(s["eval"], g["eval"], 2.0e-6)
    delta = scenes.lamp(abi.LIGHT_DISTANT, scenes.frame_z(axis), angle=0.0)
    for r in range(3):
        delta.transform[8 + r] = float(np.float32(axis[r]))
    d = ob.sample_lamp(delta, [0, 0, 0], [0, 0, 1], False, g["delta_random"])
    assert np.allclose(d["direction"], np.float32(axis), rtol=0, atol=2e-7)
    assert d["pdf"] == 1.0 and d["eval"] == 1.0


def test_delta_point_measures(golden):
    """Zero radius: pdf is the squared distance and a normalised lamp has
    eval factor 1 / (4 pi)."""
    g = golden["analytic_lamps"]["finite"]
    light = scenes.lamp(abi.LIGHT_POINT, scenes.translation(*g["center"]), size=0.0)
    s = ob.sample_lamp(light, g["reference"], g["shading_normal"], False, g["random"])
    d2 = float(np.sum((np.asarray(g["center"], np.float64) - np.asarray(g["reference"], np.float64)) ** 2))
    assert close(s["pdf"], d2, 1e-6)
    assert close(s["eval"], golden["analytic_lamps"]["delta_point_eval"], 2e-7)
    assert s["position"] == pytest.approx(g["center"], abs=1e-7)


# ---------------------------------------------------------------- renders

def plane_scene(lights, albedo=0.5):
    """A 40 x 40 diffuse plane at y = 0 seen from straight above."""
    mat = scenes.diffuse_material((albedo, albedo, albedo))
    p, t = scenes.quad([[-20, 0, -20], [-20, 0, 20], [20, 0, 20], [20, 0, -20]])
    mesh = scenes.MeshData(p, t, [0], name="plane")
    inst = scenes.InstanceData(0, scenes.identity(), name="plane")
    cam_m = scenes.translation(0.0, 4.0, 0.0) @ scenes.frame_z((0.0, 1.0, 0.0))
    camera = scenes.CameraData(cam_m, math.radians(60.0), 1.0e-3, 100.0)
    snap = scenes.SceneSnapshot([mat], [mesh], [inst], camera, (0.0, 0.0, 0.0), name="plane")
    snap.lights = lights
    return snap


def direct_settings(w=48, h=48, nee_only=True):
    s = abi.default_settings(w, h)
    s.seed = 0
    i = s.integrator
    i.max_bounces = 1
    i.min_bounces = 1
    i.transparent_max_bounces = 8
    i.light_sampling_threshold = 0.0
    if nee_only:
        i.direct_light_sampling = abi.DIRECT_LIGHT_NEE
    return s


def hit_points(snap, s):
    """World positions of the sample-0 primary hits (pixel centres)."""
    osc = ob.OracleScene(snap)
    rec = osc.primary_hits(s, 0, 1)
    t = rec[..., 2].view(np.float32).astype(np.float64)
    h, w = t.shape
    cam = np.asarray(snap.camera.transform, np.float64)
    tan_v = math.tan(0.5 * snap.camera.field_of_view)
    aspect = w / h
    ys, xs = np.mgrid[0:h, 0:w]
    sx = 2.0 * (xs + 0.5) / w - 1.0
    sy = 1.0 - 2.0 * (ys + 0.5) / h
    local = np.stack([sx * tan_v * aspect, sy * tan_v, -np.ones_like(sx)], -1)
    local /= np.linalg.norm(local, axis=-1, keepdims=True)
    d = local @ cam[:3, :3].T
    near = snap.camera.near_clip / -local[..., 2]
    o = cam[:3, 3] + d * near[..., None]
    return o + d * t[..., None], rec[..., 3] != 0


def test_delta_point_lamp_direct_closed_form():
    """L = albedo/pi * cos(theta) * power / (4 pi) / d^2 at every pixel."""
    pos = (0.3, 2.5, -0.4)
    power = 30.0
    snap = plane_scene([scenes.lamp(abi.LIGHT_POINT, scenes.translation(*pos), power=power, size=0.0)])
    s = direct_settings()
    film, _ = ob.OracleScene(snap).render(s, 0, 1, 1)
    p, valid = hit_points(snap, s)
    assert valid.all()
    off = np.asarray(pos) - p
    d2 = np.sum(off * off, -1)
    cos = off[..., 1] / np.sqrt(d2)
    expected = 0.5 / math.pi * cos * power / (4.0 * math.pi) / d2
    got = film.combined[..., 0]
    assert np.allclose(got, expected, rtol=2e-4, atol=1e-7)


def test_zero_angle_sun_direct_closed_form():
    """A zero-angle sun: L = albedo/pi * cos(theta) * power everywhere."""
    direction = np.array([0.3, 1.0, 0.2])
    snap = plane_scene([scenes.lamp(abi.LIGHT_DISTANT, scenes.frame_z(direction), power=3.0, angle=0.0)])
    s = direct_settings()
    film, _ = ob.OracleScene(snap).render(s, 0, 1, 1)
    cos = direction[1] / np.linalg.norm(direction)
    expected = 0.5 / math.pi * cos * 3.0
    assert np.allclose(film.combined[..., 0], expected, rtol=2e-5)


def test_delta_spot_cone_cut():
    """A delta spot aimed straight down: pixels outside its cone are black,
    pixels well inside it equal the point-lamp closed form."""
    pos = (0.0, 2.0, 0.0)
    spot = scenes.lamp(abi.LIGHT_SPOT, scenes.translation(*pos) @ scenes.frame_z((0.0, 1.0, 0.0)), power=30.0,
                       size=0.0, spot_angle=0.8, spot_smooth=0.0)
    snap = plane_scene([spot])
    s = direct_settings()
    film, _ = ob.OracleScene(snap).render(s, 0, 1, 1)
    p, _ = hit_points(snap, s)
    off = np.asarray(pos) - p
    d2 = np.sum(off * off, -1)
    cos = off[..., 1] / np.sqrt(d2)
    got = film.combined[..., 0]
    half = 0.4
    outside = cos < math.cos(half) - 1e-3
    inside = cos > math.cos(half) + 1e-3
    assert outside.any() and inside.any()
    assert np.all(got[outside] == 0.0)
    expected = 0.5 / math.pi * cos * 30.0 / (4.0 * math.pi) / d2
    assert np.allclose(got[inside], expected[inside], rtol=2e-4)


def test_sphere_lamp_mis_forward_nee_agree():
    """A finite sphere lamp lit by forward hits only, NEE only and MIS: the
    three estimators converge to the same image mean."""
    snap = plane_scene([scenes.lamp(abi.LIGHT_POINT, scenes.translation(0.0, 1.2, 0.0), power=40.0, size=0.35)])
    means = {}
    for name, mode in (("mis", abi.DIRECT_LIGHT_MIS), ("nee", abi.DIRECT_LIGHT_NEE),
                       ("forward", abi.DIRECT_LIGHT_FORWARD)):
        s = direct_settings(32, 32, nee_only=False)
        s.integrator.direct_light_sampling = mode
        film, _ = ob.OracleScene(snap).render(s, 0, 256, 256)
        img = film.combined[..., 0] / film.sample_count
        means[name] = float(img.mean())
    assert means["mis"] > 0.0
    assert abs(means["nee"] - means["mis"]) < 0.02 * means["mis"], means
    assert abs(means["forward"] - means["mis"]) < 0.05 * means["mis"], means
