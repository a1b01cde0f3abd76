"""Area lamps and the orthographic camera in the CPU oracle.

Pinned by:

* the Cycles renders held by tests/test_luisa_area_light_forward.cpp:
  a 32 x 32 orthographic view of a diffuse plane under a 0.8 x 0.5
  rectangle lamp (full spread; flat CDF and light tree) and a narrow-spread
  (1.2 rad) ellipse, one sample per pixel.  Frame means and the recorded
  pixels must agree within that test's own tolerances
  (:545-556 and :890-900; the values at :1012-1155);
* closed forms: the solid angle of a rectangle seen from a point on its
  axis, 4 asin(ab / sqrt((a^2 + 4h^2)(b^2 + 4h^2))), is 1 / pdf; the
  cosine-weighted solid angle estimated from stratified samples equals pi
  times the point-to-rectangle form factor; a full-spread lamp evaluates to
  1 / (pi area);
* consistency: a direction sampled from the lamp intersects it at the
  sampled distance with the same pdf and evaluation (rectangle, clamped
  rectangle and clamped ellipse).
"""
import math

import numpy as np

import oracle_binding as ob
from psycles_amd import abi, scenes


def fixture(ellipse=False, tree=False):
    """The scene and settings of tests/test_luisa_area_light_forward.cpp:153-470, 760-790."""
    mat = scenes.diffuse_material((0.62, 0.41, 0.23), name="diffuse")
    pos = np.array([[-4, -4, 0], [4, -4, 0], [4, 4, 0], [-4, 4, 0]], dtype=np.float32)
    tris = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    mesh = scenes.MeshData(pos, tris, [0], name="plane")
    inst = scenes.InstanceData(0, scenes.identity(), name="plane")
    cam = scenes.CameraData(scenes.translation(0.0, 0.0, 3.0), 0.4710899591445923, 0.10000000149011612, 1000.0,
                            sensor_fit=2, projection=1, orthographic_scale=2.200000047683716,
                            horizontal_field_of_view=0.6911112070083618)
    snap = scenes.SceneSnapshot([mat], [mesh], [inst], cam, (0.0, 0.0, 0.0), name="area_fixture")
    l = scenes.lamp(abi.LIGHT_AREA, scenes.translation(0.3700000047683716, -0.20999999344348907, 1.399999976158142),
                    (0.36000001430511475, 0.7200000286102295, 1.0), 37.0, size=0.800000011920929,
                    sphere=False, visibility=abi.VIS_ALL & ~abi.VIS_CAMERA)
    l.size_y = 0.5
    l.spread = 1.2 if ellipse else math.pi
    if ellipse:
        l.flags |= abi.LIGHT_FLAG_ELLIPSE
    snap.lights = [l]
    s = abi.default_settings(32, 32)
    s.seed = 20903
    s.pixel_filter = abi.FILTER_BLACKMAN_HARRIS
    s.filter_width = 1.5
    i = s.integrator
    i.max_bounces, i.min_bounces, i.diffuse_bounces, i.glossy_bounces = 1, 0, 0, 0
    i.transmission_bounces, i.volume_bounces, i.transparent_min_bounces, i.transparent_max_bounces = 0, 0, 0, 8
    i.sample_clamp_direct, i.sample_clamp_indirect, i.film_exposure = 0.0, 10.0, 1.0
    i.light_sampling_threshold = 0.009999999776482582
    i.use_light_tree = 1 if tree else 0
    i.direct_light_sampling = abi.DIRECT_LIGHT_MIS
    return snap, s


RECT_MEANS = (0.21808528017572826, 0.2884353661502246, 0.2247294618464366)
RECT_PIXELS = {(20, 14): (0.37258365750312805, 0.4927719235420227, 0.38393476605415344)}
ELLIPSE_MEANS = (0.5357027209195167, 0.7085100397669066, 0.552023389617716)
ELLIPSE_PIXELS = {(0, 0): (0.0, 0.0, 0.0),
                  (12, 10): (0.07275009155273438, 0.09621785581111908, 0.07496648281812668),
                  (17, 16): (0.7915834188461304, 1.0469329357147217, 0.8156996369361877),
                  (20, 14): (2.6395480632781982, 3.4910149574279785, 2.719963788986206),
                  (24, 22): (1.5788248777389526, 2.088123083114624, 1.6269252300262451)}


def render_fixture(ellipse, tree):
    snap, s = fixture(ellipse, tree)
    film, _ = ob.OracleScene(snap).render(s, 0, 1, 1)
    assert np.all(film.sample_count == 1)
    return film.combined[..., :3].astype(np.float64)


def check_against_cycles(c, means, pixels):
    m = c.reshape(-1, 3).mean(axis=0)
    assert np.all(np.abs(m - np.asarray(means)) <= 2.0e-6), m
    for (x, y), v in pixels.items():
        for a, e in zip(c[y, x], v):
            assert abs(a - e) <= 2.0e-6 * (1.0 + max(abs(a), abs(e))), ((x, y), c[y, x], v)


def test_rectangle_matches_cycles():
    check_against_cycles(render_fixture(False, False), RECT_MEANS, RECT_PIXELS)


def test_rectangle_through_light_tree_matches_cycles():
    check_against_cycles(render_fixture(False, True), RECT_MEANS, RECT_PIXELS)


def test_narrow_spread_ellipse_matches_cycles():
    check_against_cycles(render_fixture(True, False), ELLIPSE_MEANS, ELLIPSE_PIXELS)


def area_lamp(a, b, height, spread=math.pi, ellipse=False):
    """An a x b lamp at z = height facing -z (its z axis is +z)."""
    l = scenes.lamp(abi.LIGHT_AREA, scenes.translation(0.0, 0.0, height), (1.0, 1.0, 1.0), 1.0, size=a,
                    sphere=False)
    l.size_y = b
    l.spread = spread
    if ellipse:
        l.flags |= abi.LIGHT_FLAG_ELLIPSE
    return l


def test_solid_angle_closed_form_and_full_spread_evaluation():
    for a, b, h in ((0.8, 0.5, 1.4), (2.0, 1.0, 0.3), (0.1, 3.0, 2.0)):
        l = area_lamp(a, b, h)
        omega = 4.0 * math.asin(a * b / math.sqrt((a * a + 4 * h * h) * (b * b + 4 * h * h)))
        r = ob.sample_lamp(l, (0, 0, 0), (0, 0, 1), False, (0.3, 0.7))
        assert r["valid"]
        assert abs(r["pdf"] * omega - 1.0) < 1e-5
        assert abs(r["eval"] * math.pi * a * b - 1.0) < 1e-6
    le = area_lamp(0.8, 0.5, 1.0, ellipse=True)
    r = ob.sample_lamp(le, (0, 0, 0), (0, 0, 1), False, (0.3, 0.7))
    assert abs(r["eval"] * math.pi * (math.pi * 0.8 * 0.5 / 4.0) - 1.0) < 1e-6


def corner_form_factor(x, y):
    """Point-to-rectangle form factor, the point under one corner of an x by
    y (in units of the height) parallel rectangle."""
    ax, ay = math.sqrt(1 + x * x), math.sqrt(1 + y * y)
    return (x / ax * math.atan(y / ax) + y / ay * math.atan(x / ay)) / (2.0 * math.pi)


def test_samples_are_uniform_in_solid_angle():
    a, b, h = 1.2, 0.7, 0.9
    l = area_lamp(a, b, h)
    n = 48
    total = 0.0
    for i in range(n):
        for j in range(n):
            r = ob.sample_lamp(l, (0, 0, 0), (0, 0, 1), False, ((i + 0.5) / n, (j + 0.5) / n))
            assert r["valid"]
            p = r["position"]
            assert abs(p[2] - h) < 1e-5 and abs(p[0]) <= a / 2 + 1e-5 and abs(p[1]) <= b / 2 + 1e-5
            total += r["direction"][2] / r["pdf"]
    estimate = total / (n * n)
    exact = math.pi * 4.0 * corner_form_factor(a / 2 / h, b / 2 / h)
    assert abs(estimate / exact - 1.0) < 2e-3


def test_sampled_directions_hit_the_lamp_with_the_same_pdf():
    rng = np.random.default_rng(7)
    lamps = [area_lamp(0.8, 0.5, 1.4), area_lamp(0.8, 0.5, 1.4, spread=0.9),
             area_lamp(1.0, 1.0, 1.1, spread=0.6, ellipse=True), area_lamp(0.9, 0.4, 1.0, spread=1.3, ellipse=True)]
    checked = 0
    for l in lamps:
        for _ in range(40):
            ref = (rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), 0.0)
            r = ob.sample_lamp(l, ref, (0, 0, 1), False, tuple(rng.uniform(0, 1, 2)))
            if not r["valid"]:
                continue
            h = ob.intersect_lamp(l, ref, r["direction"], 0.0, 1e30, (0, 0, 1), False)
            assert h["valid"]
            assert abs(h["distance"] - r["distance"]) <= 1e-5 * r["distance"]
            assert abs(h["pdf"] - r["pdf"]) <= 2e-4 * r["pdf"], (h, r)
            assert abs(h["eval"] - r["eval"]) <= 2e-4 * r["eval"]
            checked += 1
    assert checked > 100
