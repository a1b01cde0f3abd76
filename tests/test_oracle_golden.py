"""Pin the CPU oracle against the reference's own known-answer vectors.

Every expectation here comes from a reference test file (cited in
tests/golden/reference_vectors.json) or from the reference's sampling sources
compiled in place (oracle/_ref), never from the oracle itself.
"""
import ctypes as C
import math
import struct

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi


def bits(f):
    return struct.unpack("<I", struct.pack("<f", f))[0]


def hexbits(values):
    return [int(v, 16) for v in values]


# ---------------------------------------------------------------- Sobol LUT

def fnv1a64(table: np.ndarray) -> int:
    data = table.astype("<f4").tobytes()
    h = 14695981039346656037
    # vectorised FNV-1a is sequential; process in Python over bytes (1 MiB)
    prime = 1099511628211
    mask = (1 << 64) - 1
    for byte in data:
        h ^= byte
        h = (h * prime) & mask
    return h


def test_sequence_size_for_samples(golden):
    lib = ob.load()
    for count, expected in golden["tabulated_sobol"]["sequence_size_for_samples"]:
        assert lib.oracle_sequence_size_for_samples(count) == expected


def test_sobol_table_rows(golden):
    t = ob.sobol_table(256)
    assert t.shape == (256 * 256, 4)
    assert [bits(v) for v in t[0]] == hexbits(golden["tabulated_sobol"]["row0_bits"])
    assert [bits(v) for v in t[1]] == hexbits(golden["tabulated_sobol"]["row1_bits"])


def test_sobol_table_fingerprint(golden):
    t = ob.sobol_table(256)
    assert str(fnv1a64(t)) == golden["tabulated_sobol"]["table_256_fnv1a64"]


def test_sobol_table_matches_reference_sources():
    ref = ob.ref_library()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for size in (256, 512, 4096):
        mine = ob.sobol_table(size)
        theirs = np.zeros_like(mine)
        assert ref.ref_sobol_table(size, ob.fptr(theirs)) == 0
        assert np.array_equal(mine.view(np.uint32), theirs.view(np.uint32))
    for count in (0, 1, 7, 255, 256, 257, 1000, 4096, 9000):
        assert ob.load().oracle_sequence_size_for_samples(count) == \
            ref.ref_sequence_size_for_samples(count)


def test_sobol_strata_invariant():
    """Progressive stratification: the first 2^k points of every pattern put
    exactly one point in each of the 2^k strata of every dimension."""
    t = ob.sobol_table(256).reshape(256, 256, 4)
    for k in (1, 4, 8):
        n = 1 << k
        strata = np.floor(t[:, :n, :] * n).astype(np.int64)
        for pattern in range(0, 256, 17):
            for dim in range(4):
                assert sorted(strata[pattern, :, dim].tolist()) == list(range(n))


# ---------------------------------------------------------------- sampler

def test_device_sampler_fixture(golden):
    g = golden["device_sampler"]
    lib = ob.load()
    x, y = g["pixel"]
    rng_hash = lib.oracle_pixel_hash(x, y, g["seed"])
    assert rng_hash == int(g["rng_hash"], 16)
    table = ob.sobol_table(256)
    for probe in g["probes"]:
        index = lib.oracle_shuffled_sample_index(probe["sample"], probe["dimension"], rng_hash, 256)
        assert [bits(v) for v in table[index]] == hexbits(probe["bits"])
    for entry in g["shuffled_indices"]:
        assert lib.oracle_shuffled_sample_index(entry["sample"], entry["dimension"],
                                                rng_hash, 256) == entry["index"]


def test_sampler_probe_layout_matches_fixture(golden):
    """The whole-window probe used by the GPU parity tests reproduces the same
    bits at pixel (17, 29): the fixture's (x, y) is already in Cycles' lower-left
    row convention, so the window row is H - 1 - 29."""
    g = golden["device_sampler"]
    s = abi.default_settings(32, 40)
    out = ob.sampler_probe(s, 63, 256, 17, 1)
    row = 40 - 1 - g["pixel"][1]
    rec = out[row, g["pixel"][0]]
    assert rec[0] == int(g["rng_hash"], 16)
    assert list(rec[1:5]) == hexbits(g["probes"][1]["bits"])


# ---------------------------------------------------------------- pixel filter

def check_filter_table(t, lo, hi, tol):
    half = 511
    assert np.all(np.isfinite(t))
    assert abs(t[half] - 0.5) <= tol
    assert abs(t[0] - lo) <= tol
    assert abs(t[1022] - hi) <= tol
    assert t[1023] == 0.0
    for i in range(half + 1):
        assert abs(t[half - i] + t[half + i] - 1.0) <= tol
        if i:
            assert t[half - i] <= t[half - i + 1]
            assert t[half + i] >= t[half + i - 1]


def test_pixel_filter_tables(golden):
    g = golden["pixel_filter"]
    tol = g["tolerance"]
    box = ob.filter_table(abi.FILTER_BOX, g["box_width"])
    check_filter_table(box, *g["box_endpoints"], tol)
    assert abs(box[510] - g["box_510"]) <= tol
    assert abs(box[512] - g["box_512"]) <= tol
    check_filter_table(ob.filter_table(abi.FILTER_GAUSSIAN, g["gaussian_width"]),
                       *g["gaussian_endpoints"], tol)
    check_filter_table(ob.filter_table(abi.FILTER_BLACKMAN_HARRIS, g["blackman_harris_width"]),
                       *g["blackman_harris_endpoints"], tol)


def test_pixel_filter_matches_reference_sources():
    ref = ob.ref_library()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for kind, width in ((0, 1.0), (0, 2.5), (1, 1.5), (1, 2.25), (2, 1.5), (2, 3.0)):
        mine = ob.filter_table(kind, width)
        theirs = np.zeros(1024, dtype=np.float32)
        ref.ref_pixel_filter_table(kind, width, ob.fptr(theirs))
        assert np.array_equal(mine.view(np.uint32), theirs.view(np.uint32)), (kind, width)


def test_pixel_filter_lookup(golden):
    g = golden["camera"]
    table = ob.filter_table(abi.FILTER_BLACKMAN_HARRIS, g["filter_table"]["width"])
    lib = ob.load()
    for u in g["filter_samples"]:
        coord = min(max(u, 0.0), 1.0) * 1023.0
        i = min(int(coord), 1023)
        n = min(i + 1, 1023)
        t = coord - i
        expected = table[i] * (1.0 - t) + table[n] * t
        assert abs(lib.oracle_pixel_filter_sample(ob.fptr(table), u) - expected) <= g["filter_tolerance"]


# ---------------------------------------------------------------- mappings

def test_uniform_disk(golden):
    g = golden["uniform_disk"]
    lib = ob.load()
    for r, e in zip(g["random"], g["expected"]):
        ri = np.asarray(r, dtype=np.float32)
        out = np.zeros(2, dtype=np.float32)
        lib.oracle_sample_uniform_disk(ob.fptr(ri), ob.fptr(out))
        assert np.all(np.abs(out - e) <= g["tolerance"])


def test_cosine_hemisphere(golden):
    g = golden["cosine_hemisphere"]
    lib = ob.load()
    tol = g["tolerance"]
    for case in g["cases"]:
        n = np.asarray(case["normal"], dtype=np.float32)
        r = np.asarray(case["random"], dtype=np.float32)
        out = np.zeros(4, dtype=np.float32)
        lib.oracle_sample_cosine_hemisphere(ob.fptr(n), ob.fptr(r), ob.fptr(out))
        for a, e in zip(out[:3], case["direction"]):
            assert abs(a - e) <= tol * max(1.0, abs(a), abs(e))
        assert abs(out[3] - case["pdf"]) <= tol * max(1.0, abs(case["pdf"]))


def test_cosine_hemisphere_is_cosine_distributed():
    """Closed form: E[cos] = 2/3 and the returned pdf equals cos/pi."""
    lib = ob.load()
    rng = np.random.default_rng(7)
    n = np.asarray([0.0, 0.0, 1.0], dtype=np.float32)
    cos = []
    for _ in range(4000):
        r = rng.random(2).astype(np.float32)
        out = np.zeros(4, dtype=np.float32)
        lib.oracle_sample_cosine_hemisphere(ob.fptr(n), ob.fptr(r), ob.fptr(out))
        assert abs(np.linalg.norm(out[:3]) - 1.0) < 1e-5
        assert abs(out[3] - out[2] / math.pi) < 1e-6
        cos.append(out[2])
    assert abs(np.mean(cos) - 2.0 / 3.0) < 0.01


# ---------------------------------------------------------------- camera

def test_camera_clip_range(golden):
    lib = ob.load()
    for case in golden["camera"]["clip_cases"]:
        out = np.zeros(2, dtype=np.float32)
        lib.oracle_camera_clip_range(case["near"], case["far"], case["cosine"], ob.fptr(out))
        dp, dd = case["differential"]
        clipped_dp = np.float32(dp) + out[0] * np.float32(dd)
        e = case["expected"]
        assert abs(out[0] - e[0]) <= 2e-5
        assert abs(clipped_dp - e[1]) <= 2e-5
        assert abs(out[1] - e[2]) <= 5e-4
        assert abs(out[0] + out[1] - e[3]) <= 5e-4


def test_camera_affine_fma_bits(golden):
    g = golden["camera"]
    m = np.asarray(g["affine_transform"], dtype=np.float32)
    p = np.asarray(g["affine_point"], dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    ob.load().oracle_affine_point(ob.fptr(m), ob.fptr(p), ob.fptr(out))
    assert bits(out[0]) == int(g["affine_origin_x_bits"], 16)


# ---------------------------------------------------------------- path state

def state_after(label, runtime, limits, before=None):
    lib = ob.load()
    s0 = ob.PathState()
    if before is None:
        lib.oracle_initial_state(C.byref(s0))
    else:
        s0 = before
    lim = ob.PathLimits(*limits)
    s1 = ob.PathState()
    lib.oracle_next_surface(C.byref(s0), label, runtime, C.byref(lim), C.byref(s1))
    return s1


def test_path_state_records(golden):
    g = golden["path_state"]
    lib = ob.load()
    init = ob.PathState()
    lib.oracle_initial_state(C.byref(init))
    assert (init.flag, init.visibility, init.bounce, init.rng_offset) == (
        g["initial"]["flag"], g["initial"]["visibility"], g["initial"]["bounce"],
        g["initial"]["rng_offset"])
    d = g["diffuse"]
    s = state_after(d["label"], d["runtime"], d["limits"])
    assert (s.flag, s.visibility, s.bounce, s.rng_offset) == (
        d["expected_flag"], d["visibility"], d["bounce"], d["rng_offset"])
    assert [s.diffuse_bounce, s.glossy_bounce, s.transmission_bounce,
            s.transparent_bounce] == d["counters"]
    assert lib.oracle_contract_visibility(s.visibility) == d["contract_visibility"]
    assert lib.oracle_contract_visibility(2 | 4 | 8) == d["contract_visibility_transmit_diffuse_glossy"]
    t = g["transparent"]
    st = state_after(t["label"], t["runtime"], t["limits"])
    assert (st.flag, st.visibility, st.bounce, st.transparent_bounce, st.rng_offset) == (
        t["flag"], t["visibility"], t["bounce"], t["transparent_bounce"], t["rng_offset"])
    assert lib.oracle_contract_visibility(st.visibility) == t["contract_visibility"]
    sg = g["singular"]
    s = state_after(sg["label"], sg["runtime"], sg["limits"])
    assert (s.flag, s.visibility, s.glossy_bounce, s.rng_offset) == (
        sg["expected_flag"], sg["visibility"], sg["glossy_bounce"], sg["rng_offset"])
    tt = g["transparent_twice"]
    s2 = state_after(t["label"], t["runtime"], tt["limits"], before=st)
    assert (s2.rng_offset, s2.rng_offset + 1, s2.bounce, s2.transparent_bounce) == (
        tt["rng_offset"], tt["light_dimension"], tt["bounce"], tt["transparent_bounce"])
    for case in g["continuation_probability"]:
        th = np.asarray(case["throughput"], dtype=np.float32)
        p = lib.oracle_continuation_probability(case["flag"], case["bounce"],
                                                case["transparent_bounce"], case["minimum"],
                                                case["transparent_minimum"], ob.fptr(th))
        assert p == pytest.approx(case["expected"], abs=1e-7)


# ---------------------------------------------------------------- light CDF

def distribution(areas, analytic, env):
    lib = ob.load()
    a = np.asarray(areas, dtype=np.float32)
    entries = np.zeros((len(areas) + analytic + 2, 4), dtype=np.uint32)
    tp = C.c_float()
    lp = C.c_float()
    n = lib.oracle_light_distribution(ob.fptr(a) if a.size else None, len(areas), analytic,
                                      1 if env else 0, ob.uptr(entries), C.byref(tp),
                                      C.byref(lp))
    cdf = entries[: n + 1, 0].view(np.float32)
    pdf = entries[:n, 1].view(np.float32)
    return n, cdf, pdf, entries[:n, 2], entries[:n, 3], tp.value, lp.value


def test_light_distribution(golden):
    g = golden["light_distribution"]
    tol = g["tolerance"]
    m = g["mixed"]
    n, cdf, pdf, kinds, idx, tp, lp = distribution(m["areas"], m["analytic"], m["environment"])
    assert n == m["emitter_count"]
    assert abs(tp - m["triangle_area_pdf"]) <= tol
    assert abs(lp - m["light_selection_pdf"]) <= tol
    assert np.all(np.abs(cdf - m["cdf"]) <= tol)
    assert np.all(np.abs(pdf - m["pdf"]) <= tol)
    assert list(kinds) == m["kinds"] and list(idx) == m["indices"]
    dg = g["degenerate"]
    n, cdf, pdf, *_ = distribution(dg["areas"], 0, False)
    assert np.all(np.abs(pdf - dg["pdf"]) <= tol)
    lo = g["lights_only"]
    n, cdf, pdf, *_ = distribution([], lo["analytic"], False)
    assert np.all(np.abs(pdf - lo["pdf"]) <= tol)


def test_light_distribution_matches_reference_sources():
    ref = ob.ref_library()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(3)
    for trial in range(20):
        count = int(rng.integers(0, 40))
        areas = (rng.random(count) * 3.0).astype(np.float32)
        if count:
            areas[rng.random(count) < 0.1] = 0.0
        analytic = int(rng.integers(0, 5))
        env = bool(rng.integers(0, 2))
        e_mine = np.zeros((count + analytic + 2, 4), dtype=np.uint32)
        e_ref = np.zeros_like(e_mine)
        tp1, lp1, tp2, lp2 = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        aptr = ob.fptr(areas) if count else None
        n1 = ob.load().oracle_light_distribution(aptr, count, analytic, int(env),
                                                 ob.uptr(e_mine), C.byref(tp1), C.byref(lp1))
        n2 = ref.ref_light_distribution(aptr, count, analytic, int(env), ob.uptr(e_ref),
                                        C.byref(tp2), C.byref(lp2))
        assert n1 == n2
        assert np.array_equal(e_mine[: n1 + 1], e_ref[: n2 + 1])
        assert bits(tp1.value) == bits(tp2.value) and bits(lp1.value) == bits(lp2.value)


# ---------------------------------------------------------------- triangle test

def test_pluecker_closed_form():
    """A ray through a triangle's interior hits at the analytic distance with
    barycentrics that reproduce the hit point; a miss stays a miss; the
    closed interval accepts t == tmax and rejects t just beyond it."""
    lib = ob.load()
    p0 = np.asarray([0.0, 0.0, 0.0], dtype=np.float32)
    p1 = np.asarray([1.0, 0.0, 0.0], dtype=np.float32)
    p2 = np.asarray([0.0, 1.0, 0.0], dtype=np.float32)
    o = np.asarray([0.25, 0.25, 2.0], dtype=np.float32)
    d = np.asarray([0.0, 0.0, -1.0], dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    assert lib.oracle_pluecker_intersect(ob.fptr(o), ob.fptr(d), 0.0, 10.0, ob.fptr(p0),
                                         ob.fptr(p1), ob.fptr(p2), ob.fptr(out))
    assert out[0] == pytest.approx(2.0, abs=1e-6)
    assert out[1] == pytest.approx(0.25, abs=1e-6) and out[2] == pytest.approx(0.25, abs=1e-6)
    assert lib.oracle_pluecker_intersect(ob.fptr(o), ob.fptr(d), 0.0, float(out[0]),
                                         ob.fptr(p0), ob.fptr(p1), ob.fptr(p2), ob.fptr(out))
    below = np.nextafter(np.float32(2.0), np.float32(0.0))
    assert not lib.oracle_pluecker_intersect(ob.fptr(o), ob.fptr(d), 0.0, float(below),
                                             ob.fptr(p0), ob.fptr(p1), ob.fptr(p2),
                                             ob.fptr(out))
    o2 = np.asarray([0.8, 0.8, 2.0], dtype=np.float32)
    assert not lib.oracle_pluecker_intersect(ob.fptr(o2), ob.fptr(d), 0.0, 10.0, ob.fptr(p0),
                                             ob.fptr(p1), ob.fptr(p2), ob.fptr(out))
